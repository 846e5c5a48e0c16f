// Synthetic RGB-D sequence generator: the input source for the dense path's
// parity tests and benchmarks (it stands in for the reference's OpenNI2 camera
// and TUM readers, Thirdparty/OpenNI2/src/Camera.cpp:42-51,
// Examples/tum_rgbd.cpp:43-57).
//
// Scene: an axis-aligned box room x in [-2,2], y in [0,3] (y up),
// z in [-1.5,1.5] with three spheres (radius U(0.3,0.6), seeded centres kept
// clear of the camera path).  Walls and spheres carry an albedo textured with
// Perlin's improved noise (Perlin, SIGGRAPH 2002) and are lit by a fixed point
// light, so the colour of a surface point does not depend on the viewpoint
// (photometric constancy for the tracker).
//
// Camera: on a circle of radius `radius` around the room's vertical axis at
// height `cam_height`, looking at (0, cam_height, 0); camera frame x-right, y-down,
// z-forward; angle theta_t = theta0 + step * t.  Depth is the z-depth of the
// first hit plus N(0, (noise_k z^2)^2) noise, with Bernoulli(dropout) pixels
// zeroed, stored as uint16 = round(depth_scale * z).  Random numbers come from
// a counter-based Philox-4x32-10 generator (Salmon et al., SC 2011), keyed by
// the seed and counted by (pixel, frame), so every frame is reproducible on
// its own.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>

extern "C" {

typedef struct {
  int width, height;
  double fx, fy, cx, cy;
  uint32_t seed;           // scene (sphere placement, texture) seed
  uint32_t noise_seed;     // per-sequence noise stream
  double theta0;           // start angle (rad)
  double step;             // angle increment per frame (rad)
  double radius;           // camera circle radius (m)
  double cam_height;       // camera height (m)
  double noise_k;          // depth noise sigma = noise_k * z^2
  double dropout;          // probability that a pixel's depth is zeroed
  double depth_scale;      // uint16 units per metre
} yas_synth_params;

void yas_synth_default_params(yas_synth_params* p);
int yas_synth_frame(const yas_synth_params* p, int frame, uint16_t* depth,
                    uint8_t* rgb, float* clean_depth, double T_world_cam[16]);
int yas_synth_pose(const yas_synth_params* p, int frame, double T_world_cam[16]);
int yas_synth_render(const yas_synth_params* p, const double T_world_cam[16], int frame, uint16_t* depth,
                     uint8_t* rgb, float* clean_depth);
int yas_synth_spheres(const yas_synth_params* p, double out[12]);
}

namespace {

struct Philox {
  static inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
  }
  // Philox-4x32 with 10 rounds.
  static inline void gen(uint32_t k0, uint32_t k1, const uint32_t ctr[4], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    for (int r = 0; r < 10; ++r) {
      uint32_t hi0, lo0, hi1, lo1;
      mulhilo(0xD2511F53u, c0, hi0, lo0);
      mulhilo(0xCD9E8D57u, c2, hi1, lo1);
      uint32_t n0 = hi1 ^ c1 ^ k0;
      uint32_t n1 = lo1;
      uint32_t n2 = hi0 ^ c3 ^ k1;
      uint32_t n3 = lo0;
      c0 = n0; c1 = n1; c2 = n2; c3 = n3;
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
  }
};

// uniform in (0, 1]
inline double u01(uint32_t x) { return ((double)(x >> 8) + 1.0) * (1.0 / 16777216.0); }

struct Perlin {
  int p[512];
  explicit Perlin(uint32_t seed) {
    int perm[256];
    for (int i = 0; i < 256; ++i) perm[i] = i;
    for (int i = 255; i > 0; --i) {
      uint32_t ctr[4] = {(uint32_t)i, 0x7E27u, 0u, 0u}, o[4];
      Philox::gen(seed, 0x51A7u, ctr, o);
      int j = (int)(o[0] % (uint32_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    for (int i = 0; i < 512; ++i) p[i] = perm[i & 255];
  }
  static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
  static double lerp(double t, double a, double b) { return a + t * (b - a); }
  static double grad(int h, double x, double y, double z) {
    h &= 15;
    double u = h < 8 ? x : y;
    double v = h < 4 ? y : ((h == 12 || h == 14) ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
  }
  double noise(double x, double y, double z) const {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
    x -= fx; y -= fy; z -= fz;
    double u = fade(x), v = fade(y), w = fade(z);
    int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
    int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
    return lerp(w,
                lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                     lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(p[AB + 1], x, y - 1, z - 1), grad(p[BB + 1], x - 1, y - 1, z - 1))));
  }
};

struct Sphere { double c[3]; double r; double col[3]; };

struct Scene {
  Sphere s[3];
  Perlin perlin;
  explicit Scene(const yas_synth_params& P) : perlin(P.seed) {
    uint32_t attempt = 0;
    for (int i = 0; i < 3; ++i) {
      for (;;) {
        uint32_t ctr[4] = {(uint32_t)i, attempt++, 0x5348u, 0u}, o[4];
        Philox::gen(P.seed, 0xC0FFEEu, ctr, o);
        double r = 0.3 + 0.3 * u01(o[0]);
        double m = r + 0.1;
        double cx = (-2.0 + m) + (4.0 - 2 * m) * u01(o[1]);
        double cy = (0.0 + m) + (3.0 - 2 * m) * u01(o[2]);
        double cz = (-1.5 + m) + (3.0 - 2 * m) * u01(o[3]);
        // keep the camera circle clear: distance from the centre to the circle
        double rho = std::sqrt(cx * cx + cz * cz);
        double dc = std::sqrt((rho - P.radius) * (rho - P.radius) + (cy - P.cam_height) * (cy - P.cam_height));
        bool ok = dc > r + 0.3;
        for (int j = 0; j < i && ok; ++j) {
          double dx = cx - s[j].c[0], dy = cy - s[j].c[1], dz = cz - s[j].c[2];
          ok = std::sqrt(dx * dx + dy * dy + dz * dz) > r + s[j].r + 0.05;
        }
        if (!ok && attempt < 100000) continue;
        s[i].c[0] = cx; s[i].c[1] = cy; s[i].c[2] = cz; s[i].r = r;
        uint32_t ctr2[4] = {(uint32_t)i, 0xC01u, 0u, 0u}, o2[4];
        Philox::gen(P.seed, 0xC0105u, ctr2, o2);
        s[i].col[0] = 0.35 + 0.6 * u01(o2[0]);
        s[i].col[1] = 0.35 + 0.6 * u01(o2[1]);
        s[i].col[2] = 0.35 + 0.6 * u01(o2[2]);
        break;
      }
    }
  }

  // Returns ray parameter t of the first hit along o + t d (d unnormalised),
  // the surface normal and the surface id (0..5 walls, 6..8 spheres).
  double intersect(const double o[3], const double d[3], double n[3], int& id) const {
    static const double lo[3] = {-2.0, 0.0, -1.5}, hi[3] = {2.0, 3.0, 1.5};
    double tbest = 1e30;
    id = -1;
    for (int a = 0; a < 3; ++a) {
      if (d[a] == 0.0) continue;
      double wall = d[a] > 0 ? hi[a] : lo[a];
      double t = (wall - o[a]) / d[a];
      if (t > 0 && t < tbest) {
        tbest = t;
        id = 2 * a + (d[a] > 0 ? 1 : 0);
        n[0] = n[1] = n[2] = 0.0;
        n[a] = d[a] > 0 ? -1.0 : 1.0;
      }
    }
    for (int i = 0; i < 3; ++i) {
      double oc[3] = {o[0] - s[i].c[0], o[1] - s[i].c[1], o[2] - s[i].c[2]};
      double A = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double B = 2.0 * (oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2]);
      double C = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - s[i].r * s[i].r;
      double disc = B * B - 4 * A * C;
      if (disc < 0) continue;
      double sq = std::sqrt(disc);
      double t = (-B - sq) / (2 * A);
      if (t <= 1e-9) t = (-B + sq) / (2 * A);
      if (t > 1e-9 && t < tbest) {
        tbest = t;
        id = 6 + i;
        for (int k = 0; k < 3; ++k) n[k] = (o[k] + t * d[k] - s[i].c[k]) / s[i].r;
      }
    }
    return tbest;
  }

  void albedo(int id, const double p[3], double rgb[3]) const {
    static const double wall[6][3] = {{0.95, 0.85, 0.7}, {0.7, 0.9, 0.95}, {0.8, 0.8, 0.8},
                                      {0.95, 0.95, 0.9}, {0.9, 0.75, 0.8}, {0.75, 0.95, 0.75}};
    double a = perlin.noise(3.1 * p[0], 3.1 * p[1], 3.1 * p[2]) +
               0.5 * perlin.noise(7.3 * p[0] + 11.0, 7.3 * p[1], 7.3 * p[2]);
    double tex = std::min(1.0, std::max(0.08, 0.55 + 0.45 * a));
    const double* base = id < 6 ? wall[id] : s[id - 6].col;
    for (int k = 0; k < 3; ++k) rgb[k] = base[k] * tex;
  }
};

void pose_of(const yas_synth_params& P, int frame, double T[16]) {
  double th = P.theta0 + P.step * frame;
  double c[3] = {P.radius * std::cos(th), P.cam_height, P.radius * std::sin(th)};
  double f[3] = {-c[0], 0.0, -c[2]};
  double fn = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (double& v : f) v /= fn;
  // right = forward x up(0,1,0); down = forward x right
  double r[3] = {-f[2], 0.0, f[0]};
  double rn = std::sqrt(r[0] * r[0] + r[2] * r[2]);
  r[0] /= rn; r[2] /= rn;
  double dn[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]};
  for (int i = 0; i < 3; ++i) {
    T[4 * i + 0] = r[i];
    T[4 * i + 1] = dn[i];
    T[4 * i + 2] = f[i];
    T[4 * i + 3] = c[i];
  }
  T[12] = T[13] = T[14] = 0.0;
  T[15] = 1.0;
}

}  // namespace

extern "C" {

void yas_synth_default_params(yas_synth_params* p) {
  p->width = 640;
  p->height = 480;
  p->fx = p->fy = 525.0;
  p->cx = 319.5;
  p->cy = 239.5;
  p->seed = 0;
  p->noise_seed = 0;
  p->theta0 = 0.0;
  p->step = 0.02;
  p->radius = 1.0;
  p->cam_height = 1.5;
  p->noise_k = 0.0012;
  p->dropout = 0.02;
  p->depth_scale = 1000.0;
}

int yas_synth_pose(const yas_synth_params* p, int frame, double T[16]) {
  if (!p || !T) return 1;
  pose_of(*p, frame, T);
  return 0;
}

int yas_synth_spheres(const yas_synth_params* p, double out[12]) {
  Scene sc(*p);
  for (int i = 0; i < 3; ++i) {
    out[4 * i + 0] = sc.s[i].c[0];
    out[4 * i + 1] = sc.s[i].c[1];
    out[4 * i + 2] = sc.s[i].c[2];
    out[4 * i + 3] = sc.s[i].r;
  }
  return 0;
}

// The scene of `p` seen from an arbitrary camera pose T (camera -> world);
// `frame` selects the noise and dropout draws.
int yas_synth_render(const yas_synth_params* p, const double T_in[16], int frame, uint16_t* depth, uint8_t* rgb,
                     float* clean_depth) {
  if (!p || !T_in || p->width <= 0 || p->height <= 0) return 1;
  const yas_synth_params& P = *p;
  Scene sc(P);
  double T[16];
  std::memcpy(T, T_in, sizeof(T));
  const double light[3] = {0.3, 2.8, -0.2};
  const int W = P.width, H = P.height;
#pragma omp parallel for schedule(static)
  for (int v = 0; v < H; ++v) {
    for (int u = 0; u < W; ++u) {
      double dc[3] = {(u - P.cx) / P.fx, (v - P.cy) / P.fy, 1.0};
      double dw[3];
      for (int i = 0; i < 3; ++i) dw[i] = T[4 * i] * dc[0] + T[4 * i + 1] * dc[1] + T[4 * i + 2] * dc[2];
      double o[3] = {T[3], T[7], T[11]};
      double n[3];
      int id;
      double z = sc.intersect(o, dw, n, id);
      const size_t k = (size_t)v * W + u;
      if (clean_depth) clean_depth[k] = (float)z;
      uint32_t ctr[4] = {(uint32_t)k, (uint32_t)frame, 0xD3Bu, 0u}, o4[4];
      Philox::gen(P.noise_seed, 0x5EEDu + P.seed, ctr, o4);
      double g = std::sqrt(-2.0 * std::log(u01(o4[0]))) * std::cos(6.283185307179586 * u01(o4[1]));
      double zn = z + g * P.noise_k * z * z;
      bool drop = u01(o4[2]) <= P.dropout;
      if (depth) {
        double q = std::nearbyint(P.depth_scale * zn);
        depth[k] = (drop || !(q > 0.0) || q > 65535.0) ? 0 : (uint16_t)q;
      }
      if (rgb) {
        double pw[3] = {o[0] + z * dw[0], o[1] + z * dw[1], o[2] + z * dw[2]};
        double alb[3];
        sc.albedo(id, pw, alb);
        double l[3] = {light[0] - pw[0], light[1] - pw[1], light[2] - pw[2]};
        double ln = std::sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
        double ndl = (n[0] * l[0] + n[1] * l[1] + n[2] * l[2]) / ln;
        double shade = 0.3 + 0.7 * std::max(0.0, ndl);
        for (int c = 0; c < 3; ++c) {
          double val = std::nearbyint(255.0 * alb[c] * shade);
          rgb[3 * k + c] = (uint8_t)std::min(255.0, std::max(0.0, val));
        }
      }
    }
  }
  return 0;
}

int yas_synth_frame(const yas_synth_params* p, int frame, uint16_t* depth, uint8_t* rgb,
                    float* clean_depth, double T_world_cam[16]) {
  if (!p) return 1;
  double T[16];
  pose_of(*p, frame, T);
  if (T_world_cam) std::memcpy(T_world_cam, T, sizeof(T));
  return yas_synth_render(p, T, frame, depth, rgb, clean_depth);
}

}  // extern "C"
