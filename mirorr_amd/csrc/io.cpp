// ITK-free image and transform file I/O of the command-line drop-in
// (include/mirorr_io.h). Host code only.
#include "../../include/mirorr_io.h"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "xform.hpp"

namespace {

thread_local std::string g_io_error;

struct IoError : std::runtime_error {
  mirorr_status code;
  IoError(mirorr_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <class F>
mirorr_status io_guard(F&& f) {
  try {
    f();
    g_io_error.clear();
    return MIRORR_OK;
  } catch (const IoError& e) {
    g_io_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_io_error = "host allocation failed";
    return MIRORR_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_io_error = e.what();
    return MIRORR_ERR_PLATFORM;
  }
}

[[noreturn]] void fail(const std::string& m) { throw IoError(MIRORR_ERR_ARG, m); }

bool ends_with(const std::string& s, const char* suf) {
  const size_t n = std::strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

// ---- NIfTI-1 header (348 bytes, little or big endian) -----------------------
constexpr int kHdrSize = 348;

struct Hdr {
  bool swap = false;
  int16_t dim[8] = {};
  int16_t datatype = 0, bitpix = 0;
  float pixdim[8] = {};
  float vox_offset = 0.0f, scl_slope = 0.0f, scl_inter = 0.0f;
  int16_t qform_code = 0, sform_code = 0;
  float quatern[3] = {}, qoffset[3] = {};
  float srow[3][4] = {};
};

template <class T>
T get(const unsigned char* b, int off, bool swap) {
  unsigned char t[sizeof(T)];
  std::memcpy(t, b + off, sizeof(T));
  if (swap)
    for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(t[i], t[sizeof(T) - 1 - i]);
  T v;
  std::memcpy(&v, t, sizeof(T));
  return v;
}

template <class T>
void put(unsigned char* b, int off, T v) {
  std::memcpy(b + off, &v, sizeof(T));
}

Hdr parse_header(const unsigned char* b) {
  Hdr h;
  const int32_t sz = get<int32_t>(b, 0, false);
  if (sz == kHdrSize) {
    h.swap = false;
  } else if (get<int32_t>(b, 0, true) == kHdrSize) {
    h.swap = true;
  } else {
    fail("not a NIfTI-1 file (sizeof_hdr is not 348)");
  }
  if (std::memcmp(b + 344, "n+1", 4) != 0 && std::memcmp(b + 344, "ni1", 4) != 0)
    fail("not a NIfTI-1 file (bad magic)");
  if (std::memcmp(b + 344, "ni1", 4) == 0)
    fail("two-file NIfTI (.hdr/.img) is not supported; use a .nii or .nii.gz file");
  for (int i = 0; i < 8; ++i) h.dim[i] = get<int16_t>(b, 40 + 2 * i, h.swap);
  h.datatype = get<int16_t>(b, 70, h.swap);
  h.bitpix = get<int16_t>(b, 72, h.swap);
  for (int i = 0; i < 8; ++i) h.pixdim[i] = get<float>(b, 76 + 4 * i, h.swap);
  h.vox_offset = get<float>(b, 108, h.swap);
  h.scl_slope = get<float>(b, 112, h.swap);
  h.scl_inter = get<float>(b, 116, h.swap);
  h.qform_code = get<int16_t>(b, 252, h.swap);
  h.sform_code = get<int16_t>(b, 254, h.swap);
  for (int i = 0; i < 3; ++i) h.quatern[i] = get<float>(b, 256 + 4 * i, h.swap);
  for (int i = 0; i < 3; ++i) h.qoffset[i] = get<float>(b, 268 + 4 * i, h.swap);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c) h.srow[r][c] = get<float>(b, 280 + 16 * r + 4 * c, h.swap);
  return h;
}

// NIfTI-1 quaternion (b, c, d; a >= 0 implied) to a rotation matrix.
void quatern_to_rot(double b, double c, double d, double R[3][3]) {
  double a = 1.0 - (b * b + c * c + d * d);
  if (a < 1.0e-7) {  // special case in the NIfTI-1 reference code
    a = 1.0 / std::sqrt(b * b + c * c + d * d);
    b *= a;
    c *= a;
    d *= a;
    a = 0.0;
  } else {
    a = std::sqrt(a);
  }
  R[0][0] = a * a + b * b - c * c - d * d;
  R[0][1] = 2.0 * (b * c - a * d);
  R[0][2] = 2.0 * (b * d + a * c);
  R[1][0] = 2.0 * (b * c + a * d);
  R[1][1] = a * a + c * c - b * b - d * d;
  R[1][2] = 2.0 * (c * d - a * b);
  R[2][0] = 2.0 * (b * d - a * c);
  R[2][1] = 2.0 * (c * d + a * b);
  R[2][2] = a * a + d * d - c * c - b * b;
}

// Rotation matrix (det +1) to the NIfTI quaternion (b, c, d) with a >= 0.
void rot_to_quatern(const double R[3][3], double q[3]) {
  const double tr = R[0][0] + R[1][1] + R[2][2];
  double a, b, c, d;
  if (tr > 0.0) {
    const double s = 0.5 / std::sqrt(tr + 1.0);
    a = 0.25 / s;
    b = (R[2][1] - R[1][2]) * s;
    c = (R[0][2] - R[2][0]) * s;
    d = (R[1][0] - R[0][1]) * s;
  } else if (R[0][0] > R[1][1] && R[0][0] > R[2][2]) {
    const double s = 2.0 * std::sqrt(1.0 + R[0][0] - R[1][1] - R[2][2]);
    a = (R[2][1] - R[1][2]) / s;
    b = 0.25 * s;
    c = (R[0][1] + R[1][0]) / s;
    d = (R[0][2] + R[2][0]) / s;
  } else if (R[1][1] > R[2][2]) {
    const double s = 2.0 * std::sqrt(1.0 + R[1][1] - R[0][0] - R[2][2]);
    a = (R[0][2] - R[2][0]) / s;
    b = (R[0][1] + R[1][0]) / s;
    c = 0.25 * s;
    d = (R[1][2] + R[2][1]) / s;
  } else {
    const double s = 2.0 * std::sqrt(1.0 + R[2][2] - R[0][0] - R[1][1]);
    a = (R[1][0] - R[0][1]) / s;
    b = (R[0][2] + R[2][0]) / s;
    c = (R[1][2] + R[2][1]) / s;
    d = 0.25 * s;
  }
  if (a < 0.0) {
    b = -b;
    c = -c;
    d = -d;
  }
  q[0] = b;
  q[1] = c;
  q[2] = d;
}

// LPS geometry of a header (ITK: RAS from the file, x and y negated).
mirorr_geom header_geom(const Hdr& h) {
  if (h.dim[0] < 3 || h.dim[0] > 7) fail("NIfTI dim[0] must be 3..7 for a volume");
  for (int i = 4; i <= h.dim[0]; ++i)
    if (h.dim[i] > 1) fail("only 3-D volumes are supported (dim[4..] > 1)");
  mirorr_geom g{};
  for (int i = 0; i < 3; ++i) {
    g.dims[i] = h.dim[i + 1];
    if (g.dims[i] < 1) fail("NIfTI dimensions must be positive");
    const double p = std::fabs((double)h.pixdim[i + 1]);
    g.spacing[i] = p > 0.0 ? p : 1.0;
  }
  double D[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};  // RAS direction columns
  double o[3] = {0.0, 0.0, 0.0};
  if (h.qform_code > 0) {
    quatern_to_rot(h.quatern[0], h.quatern[1], h.quatern[2], D);
    if (h.pixdim[0] < 0.0f)
      for (int r = 0; r < 3; ++r) D[r][2] = -D[r][2];
    for (int i = 0; i < 3; ++i) o[i] = h.qoffset[i];
  } else if (h.sform_code > 0) {
    for (int c = 0; c < 3; ++c) {
      double n = 0.0;
      for (int r = 0; r < 3; ++r) n += (double)h.srow[r][c] * h.srow[r][c];
      n = std::sqrt(n);
      if (!(n > 0.0)) fail("degenerate sform");
      for (int r = 0; r < 3; ++r) D[r][c] = h.srow[r][c] / n;
    }
    for (int i = 0; i < 3; ++i) o[i] = h.srow[i][3];
  }
  for (int r = 0; r < 3; ++r) {
    const double s = r < 2 ? -1.0 : 1.0;  // RAS -> LPS
    for (int c = 0; c < 3; ++c) g.direction[3 * r + c] = s * D[r][c];
    g.origin[r] = s * o[r];
  }
  return g;
}

struct Gz {
  gzFile f = nullptr;
  explicit Gz(const char* path, const char* mode) : f(gzopen(path, mode)) {}
  ~Gz() {
    if (f) gzclose(f);
  }
};

void read_exact(gzFile f, void* dst, size_t n, const char* what) {
  unsigned char* p = static_cast<unsigned char*>(dst);
  while (n > 0) {
    const unsigned chunk = (unsigned)std::min<size_t>(n, 1u << 30);
    const int r = gzread(f, p, chunk);
    if (r <= 0) fail(std::string("truncated NIfTI file (") + what + ")");
    p += r;
    n -= (size_t)r;
  }
}

// Reads n bytes into a vector that grows by chunks as data arrives, so a
// header announcing more data than the file holds fails at the end of the
// file instead of allocating the announced size first.
std::vector<unsigned char> read_growing(gzFile f, size_t n, const char* what) {
  constexpr size_t kChunk = size_t(64) << 20;
  std::vector<unsigned char> out;
  while (out.size() < n) {
    const size_t k = std::min(n - out.size(), kChunk);
    const size_t at = out.size();
    out.resize(at + k);
    read_exact(f, out.data() + at, k, what);
  }
  return out;
}

// Skips n bytes (the header extensions) without holding them.
void skip_bytes(gzFile f, size_t n, const char* what) {
  unsigned char buf[4096];
  while (n > 0) {
    const size_t k = std::min(n, sizeof(buf));
    read_exact(f, buf, k, what);
    n -= k;
  }
}

// Reads the voxels as double (scaled), plus the geometry.
std::vector<double> read_voxels(const char* path, mirorr_geom* geom) {
  if (!path) fail("null path");
  Gz gz(path, "rb");
  if (!gz.f) fail(std::string("cannot open ") + path);
  unsigned char hb[kHdrSize];
  read_exact(gz.f, hb, kHdrSize, "header");
  const Hdr h = parse_header(hb);
  *geom = header_geom(h);
  const size_t n = (size_t)geom->dims[0] * geom->dims[1] * geom->dims[2];
  int bytes = 0;
  switch (h.datatype) {
    case 2: case 256: bytes = 1; break;             // uint8, int8
    case 4: case 512: bytes = 2; break;             // int16, uint16
    case 8: case 768: case 16: bytes = 4; break;    // int32, uint32, float32
    case 64: bytes = 8; break;                      // float64
    default: fail("unsupported NIfTI datatype " + std::to_string(h.datatype));
  }
  // vox_offset is a float in the header: reject NaN, infinities and
  // offsets past any real extension block before converting it.
  if (!std::isfinite(h.vox_offset) || h.vox_offset < (double)kHdrSize ||
      h.vox_offset > double(1u << 30))
    fail("NIfTI vox_offset out of range");
  const size_t skip = (size_t)h.vox_offset - kHdrSize;
  skip_bytes(gz.f, skip, "extensions");
  const std::vector<unsigned char> raw = read_growing(gz.f, n * (size_t)bytes, "voxels");
  std::vector<double> v(n);
  const unsigned char* r = raw.data();
  for (size_t i = 0; i < n; ++i) {
    const int o = 0;
    const unsigned char* p = r + i * (size_t)bytes;
    double x = 0.0;
    switch (h.datatype) {
      case 2: x = *p; break;
      case 256: x = (int8_t)*p; break;
      case 4: x = get<int16_t>(p, o, h.swap); break;
      case 512: x = get<uint16_t>(p, o, h.swap); break;
      case 8: x = get<int32_t>(p, o, h.swap); break;
      case 768: x = get<uint32_t>(p, o, h.swap); break;
      case 16: x = get<float>(p, o, h.swap); break;
      case 64: x = get<double>(p, o, h.swap); break;
    }
    v[i] = x;
  }
  if (h.scl_slope != 0.0f && std::isfinite(h.scl_slope) &&
      !(h.scl_slope == 1.0f && h.scl_inter == 0.0f))
    for (double& x : v) x = x * h.scl_slope + h.scl_inter;
  return v;
}

void col_to_phys(const mirorr_geom& g, const double idx[3], double out[3]) {
  for (int r = 0; r < 3; ++r) {
    double s = g.origin[r];
    for (int c = 0; c < 3; ++c) s += g.direction[3 * r + c] * g.spacing[c] * idx[c];
    out[r] = s;
  }
}

void geom_center(const mirorr_geom& g, double c[3]) {
  const double idx[3] = {(g.dims[0] - 1) / 2.0, (g.dims[1] - 1) / 2.0, (g.dims[2] - 1) / 2.0};
  col_to_phys(g, idx, c);
}

void check_geom(const mirorr_geom* g) {
  if (!g) fail("null geometry");
  for (int i = 0; i < 3; ++i) {
    if (g->dims[i] < 1) fail("dimensions must be positive");
    if (!(g->spacing[i] > 0.0)) fail("spacing must be positive");
  }
}

// ---- ITK transform files ------------------------------------------------------
mirorr::xf::M3 euler_zyx(const double ang[3]) {  // Euler3D with ComputeZYX on
  const double cx = std::cos(ang[0]), sx = std::sin(ang[0]);
  const double cy = std::cos(ang[1]), sy = std::sin(ang[1]);
  const double cz = std::cos(ang[2]), sz = std::sin(ang[2]);
  mirorr::xf::M3 Rx{{{1, 0, 0}, {0, cx, -sx}, {0, sx, cx}}};
  mirorr::xf::M3 Ry{{{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}}};
  mirorr::xf::M3 Rz{{{cz, -sz, 0}, {sz, cz, 0}, {0, 0, 1}}};
  return mirorr::xf::mul(Rz, mirorr::xf::mul(Ry, Rx));
}

// GetInverseTransform of the reference (itkMirorrPyramidWrapper.txx:312-350):
// the inverse keeps the centre, its matrix is replaced by an orthogonal one.
// The reference forms U*V from vnl's SVD; the orthogonal polar factor U*V^T
// used here is what that step is meant to produce.
mirorr::xf::Xf reference_inverse(const mirorr::xf::Xf& x, int kind) {
  const mirorr::xf::Xf r = mirorr::xf::inverse(x);
  mirorr::xf::Xf o = mirorr::xf::identity(kind);
  mirorr::xf::set_center(o, r.c);
  mirorr::xf::set_translation(o, r.t);
  if (!mirorr::xf::set_matrix(o, mirorr::xf::orthogonalize(r.M)))
    throw IoError(MIRORR_ERR_ALGO, "inverse transform is not a rotation");
  return o;
}

std::vector<double> parse_numbers(const std::string& s) {
  std::istringstream in(s);
  std::vector<double> v;
  std::string tok;
  while (in >> tok) {
    char* end = nullptr;
    const double x = std::strtod(tok.c_str(), &end);
    if (end == tok.c_str() || *end != '\0') fail("bad number '" + tok + "' in transform file");
    v.push_back(x);
  }
  return v;
}

}  // namespace

extern "C" {

const char* mirorr_io_last_error(void) { return g_io_error.c_str(); }

void mirorr_io_free(void* p) { std::free(p); }

mirorr_status mirorr_nifti_read(const char* path, mirorr_geom* geom, float** data) {
  return io_guard([&] {
    if (!geom || !data) fail("null output");
    *data = nullptr;
    const std::vector<double> v = read_voxels(path, geom);
    float* out = static_cast<float*>(std::malloc(sizeof(float) * v.size()));
    if (!out) throw std::bad_alloc();
    for (size_t i = 0; i < v.size(); ++i) out[i] = (float)v[i];
    *data = out;
  });
}

mirorr_status mirorr_nifti_read_mask(const char* path, mirorr_geom* geom, uint8_t** data) {
  return io_guard([&] {
    if (!geom || !data) fail("null output");
    *data = nullptr;
    const std::vector<double> v = read_voxels(path, geom);
    uint8_t* out = static_cast<uint8_t*>(std::malloc(v.size()));
    if (!out) throw std::bad_alloc();
    // ITK's pixel conversion into an unsigned char image: a C cast, which
    // truncates fractions and keeps the low 8 bits of integers.
    for (size_t i = 0; i < v.size(); ++i) {
      const double t = std::trunc(v[i]);
      out[i] = std::isfinite(t) && std::fabs(t) < 9.0e18 ? (uint8_t)((int64_t)t & 0xFF) : 0;
    }
    *data = out;
  });
}

mirorr_status mirorr_nifti_write(const char* path, const mirorr_geom* geom, const float* data) {
  return io_guard([&] {
    if (!path || !data) fail("null argument");
    check_geom(geom);
    for (int i = 0; i < 3; ++i)
      if (geom->dims[i] > 32767) fail("NIfTI-1 dimensions are limited to 32767");
    unsigned char h[kHdrSize + 4];
    std::memset(h, 0, sizeof(h));
    put<int32_t>(h, 0, kHdrSize);
    const int16_t dim[8] = {3, (int16_t)geom->dims[0], (int16_t)geom->dims[1],
                            (int16_t)geom->dims[2], 1, 1, 1, 1};
    for (int i = 0; i < 8; ++i) put<int16_t>(h, 40 + 2 * i, dim[i]);
    put<int16_t>(h, 70, 16);  // float32
    put<int16_t>(h, 72, 32);
    // RAS direction (x and y rows negated) and its proper-rotation part.
    double D[3][3], o[3];
    for (int r = 0; r < 3; ++r) {
      const double s = r < 2 ? -1.0 : 1.0;
      for (int c = 0; c < 3; ++c) D[r][c] = s * geom->direction[3 * r + c];
      o[r] = s * geom->origin[r];
    }
    const double det = D[0][0] * (D[1][1] * D[2][2] - D[1][2] * D[2][1]) -
                       D[0][1] * (D[1][0] * D[2][2] - D[1][2] * D[2][0]) +
                       D[0][2] * (D[1][0] * D[2][1] - D[1][1] * D[2][0]);
    const float qfac = det < 0.0 ? -1.0f : 1.0f;
    double R[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) R[r][c] = (c == 2 ? qfac : 1.0) * D[r][c];
    double q[3];
    rot_to_quatern(R, q);
    const float pixdim[8] = {qfac, (float)geom->spacing[0], (float)geom->spacing[1],
                             (float)geom->spacing[2], 1, 1, 1, 1};
    for (int i = 0; i < 8; ++i) put<float>(h, 76 + 4 * i, pixdim[i]);
    put<float>(h, 108, (float)(kHdrSize + 4));  // vox_offset
    put<float>(h, 112, 1.0f);                    // scl_slope
    put<char>(h, 123, 10);                       // xyzt_units: mm, s
    put<int16_t>(h, 252, 1);                     // qform_code: scanner
    put<int16_t>(h, 254, 1);                     // sform_code: scanner
    for (int i = 0; i < 3; ++i) put<float>(h, 256 + 4 * i, (float)q[i]);
    for (int i = 0; i < 3; ++i) put<float>(h, 268 + 4 * i, (float)o[i]);
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c)
        put<float>(h, 280 + 16 * r + 4 * c, (float)(D[r][c] * geom->spacing[c]));
      put<float>(h, 280 + 16 * r + 12, (float)o[r]);
    }
    std::memcpy(h + 344, "n+1\0", 4);
    const size_t n = (size_t)geom->dims[0] * geom->dims[1] * geom->dims[2];
    const std::string p(path);
    if (ends_with(p, ".gz")) {
      Gz gz(path, "wb6");
      if (!gz.f) fail("cannot create " + p);
      if (gzwrite(gz.f, h, sizeof(h)) != (int)sizeof(h)) fail("write failed: " + p);
      const unsigned char* src = reinterpret_cast<const unsigned char*>(data);
      size_t left = n * sizeof(float);
      while (left > 0) {
        const unsigned chunk = (unsigned)std::min<size_t>(left, 1u << 30);
        if (gzwrite(gz.f, src, chunk) != (int)chunk) fail("write failed: " + p);
        src += chunk;
        left -= chunk;
      }
    } else {
      std::FILE* f = std::fopen(path, "wb");
      if (!f) fail("cannot create " + p);
      const bool ok = std::fwrite(h, 1, sizeof(h), f) == sizeof(h) &&
                      std::fwrite(data, sizeof(float), n, f) == n;
      if (std::fclose(f) != 0 || !ok) fail("write failed: " + p);
    }
  });
}

mirorr_status mirorr_reorient_ari(mirorr_geom* geom, float* data, uint8_t* mask) {
  return io_guard([&] {
    check_geom(geom);
    if (!data) fail("null data");
    const mirorr_geom g = *geom;
    // Physical axis (LPS) and sign of each image axis, column by column:
    // the strictly largest |cosine| above 0.001, ties and near-zero columns
    // going to x (ITK SpatialOrientationAdapter::FromDirectionCosines with
    // Function::Max3 and Function::Sign).
    int phys[3], sign[3];
    for (int i = 0; i < 3; ++i) {
      const double ax = std::fabs(g.direction[i]), ay = std::fabs(g.direction[3 + i]),
                   az = std::fabs(g.direction[6 + i]);
      int k = 0;
      if (ax > 0.001 && ax > ay && ax > az) k = 0;
      else if (ay > 0.001 && ay > ax && ay > az) k = 1;
      else if (az > 0.001 && az > ax && az > ay) k = 2;
      phys[i] = k;
      sign[i] = g.direction[3 * k + i] < 0.0 ? -1 : 1;
    }
    if (phys[0] == phys[1] || phys[0] == phys[2] || phys[1] == phys[2])
      fail("direction cosines too oblique to name an orientation (two image axes share a "
           "dominant physical axis)");
    // ARI: new axis 0 -> +y (anterior to posterior), 1 -> +x (right to
    // left), 2 -> +z (inferior to superior).
    const int target[3] = {1, 0, 2};
    int from[3];
    bool flip[3];
    for (int j = 0; j < 3; ++j) {
      from[j] = -1;
      for (int i = 0; i < 3; ++i)
        if (phys[i] == target[j]) from[j] = i;
      flip[j] = sign[from[j]] < 0;
    }
    mirorr_geom out{};
    double start_idx[3];
    for (int j = 0; j < 3; ++j) {
      const int i = from[j];
      out.dims[j] = g.dims[i];
      out.spacing[j] = g.spacing[i];
      for (int r = 0; r < 3; ++r)
        out.direction[3 * r + j] = (flip[j] ? -1.0 : 1.0) * g.direction[3 * r + i];
    }
    for (int i = 0; i < 3; ++i) start_idx[i] = 0.0;
    for (int j = 0; j < 3; ++j)
      if (flip[j]) start_idx[from[j]] = g.dims[from[j]] - 1;
    col_to_phys(g, start_idx, out.origin);
    const size_t n = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
    std::vector<float> tmp(data, data + n);
    std::vector<uint8_t> tmpm;
    if (mask) tmpm.assign(mask, mask + n);
    for (int k2 = 0; k2 < out.dims[2]; ++k2)
      for (int k1 = 0; k1 < out.dims[1]; ++k1)
        for (int k0 = 0; k0 < out.dims[0]; ++k0) {
          const int kn[3] = {k0, k1, k2};
          int ko[3];
          for (int j = 0; j < 3; ++j) ko[from[j]] = flip[j] ? out.dims[j] - 1 - kn[j] : kn[j];
          const size_t src = (size_t)ko[0] + (size_t)g.dims[0] * ((size_t)ko[1] + (size_t)g.dims[1] * ko[2]);
          const size_t dst = (size_t)k0 + (size_t)out.dims[0] * ((size_t)k1 + (size_t)out.dims[1] * k2);
          data[dst] = tmp[src];
          if (mask) mask[dst] = tmpm[src];
        }
    *geom = out;
  });
}

mirorr_status mirorr_centered_init(const mirorr_geom* centre_image, const mirorr_geom* other,
                                   int affine, double* params, double center[3]) {
  return io_guard([&] {
    check_geom(centre_image);
    check_geom(other);
    if (!params || !center) fail("null output");
    double c1[3], c2[3];
    geom_center(*centre_image, c1);
    geom_center(*other, c2);
    mirorr::xf::Xf x = mirorr::xf::identity(affine ? mirorr::xf::kAffine : mirorr::xf::kRigid);
    mirorr::xf::set_center(x, c1);
    const double t[3] = {c2[0] - c1[0], c2[1] - c1[1], c2[2] - c1[2]};
    mirorr::xf::set_translation(x, t);
    mirorr::xf::get_params(x, params);
    for (int i = 0; i < 3; ++i) center[i] = c1[i];
  });
}

mirorr_status mirorr_tfm_write(const char* path, int affine, const double* params,
                               const double center[3], int inverse) {
  return io_guard([&] {
    if (!path || !params || !center) fail("null argument");
    mirorr::xf::Xf x = mirorr::xf::identity(affine ? mirorr::xf::kAffine : mirorr::xf::kRigid);
    mirorr::xf::set_params(x, params);
    mirorr::xf::set_center(x, center);
    if (inverse) x = reference_inverse(x, x.kind);
    double p[12];
    mirorr::xf::get_params(x, p);
    std::ofstream f(path);
    if (!f) fail(std::string("cannot create ") + path);
    char buf[64];
    f << "#Insight Transform File V1.0\n#Transform 0\nTransform: "
      << (affine ? "AffineTransform_double_3_3" : "Euler3DTransform_double_3_3") << "\nParameters:";
    for (int i = 0; i < mirorr::xf::nparams(x); ++i) {
      std::snprintf(buf, sizeof(buf), " %.17g", p[i]);
      f << buf;
    }
    f << "\nFixedParameters:";
    for (int i = 0; i < 3; ++i) {
      std::snprintf(buf, sizeof(buf), " %.17g", x.c[i]);
      f << buf;
    }
    f << "\n";
    if (!f.good()) fail(std::string("write failed: ") + path);
  });
}

mirorr_status mirorr_tfm_read(const char* path, int want_affine, double* params,
                              double center[3], int inverse) {
  return io_guard([&] {
    if (!path || !params || !center) fail("null argument");
    std::ifstream f(path);
    if (!f) fail(std::string("cannot open ") + path);
    std::string line, cls;
    std::vector<double> p, fp;
    bool have_p = false, have_fp = false;
    while (std::getline(f, line)) {
      if (line.rfind("Transform:", 0) == 0) {
        if (!cls.empty()) break;  // the first transform only
        std::istringstream in(line.substr(10));
        in >> cls;
      } else if (line.rfind("Parameters:", 0) == 0) {
        p = parse_numbers(line.substr(11));
        have_p = true;
      } else if (line.rfind("FixedParameters:", 0) == 0) {
        fp = parse_numbers(line.substr(16));
        have_fp = true;
      }
    }
    if (cls.empty() || !have_p) fail(std::string("no transform in ") + path);
    mirorr::xf::M3 M = mirorr::xf::eye3();
    double t[3], c[3] = {0.0, 0.0, 0.0};
    if (have_fp) {
      if (fp.size() < 3) fail("transform fixed parameters must hold a centre");
      for (int i = 0; i < 3; ++i) c[i] = fp[i];
    }
    const bool euler = cls.find("Euler3DTransform") == 0;
    if (euler) {
      if (p.size() != 6) fail("Euler3DTransform needs 6 parameters");
      const double ang[3] = {p[0], p[1], p[2]};
      const bool zyx = fp.size() >= 4 && fp[3] != 0.0;
      M = zyx ? euler_zyx(ang) : mirorr::xf::euler_zxy(ang);
      for (int i = 0; i < 3; ++i) t[i] = p[3 + i];
      // The reference rebuilds the third row of an Euler file's matrix
      // as the cross product of the first two.
      M.a[2][0] = M.a[0][1] * M.a[1][2] - M.a[0][2] * M.a[1][1];
      M.a[2][1] = M.a[0][2] * M.a[1][0] - M.a[0][0] * M.a[1][2];
      M.a[2][2] = M.a[0][0] * M.a[1][1] - M.a[0][1] * M.a[1][0];
    } else if (cls.find("AffineTransform") == 0) {
      if (p.size() != 12) fail(cls + " needs 12 parameters");
      for (int i = 0; i < 9; ++i) M.a[i / 3][i % 3] = p[i];
      for (int i = 0; i < 3; ++i) t[i] = p[9 + i];
    } else {
      throw IoError(MIRORR_ERR_ARG, "transform file '" + std::string(path) +
                                        "' does not contain a MatrixOffsetTransform type (" +
                                        cls + ")");
    }
    if (cls.find("_double_3_3") == std::string::npos && cls.find("_float_3_3") == std::string::npos)
      fail("only 3-D transforms are supported (" + cls + ")");
    const int kind = want_affine ? mirorr::xf::kAffine : mirorr::xf::kRigid;
    mirorr::xf::Xf src = mirorr::xf::identity(mirorr::xf::kAffine);
    mirorr::xf::set_center(src, c);
    mirorr::xf::set_matrix(src, M);
    mirorr::xf::set_translation(src, t);
    if (inverse) src = reference_inverse(src, kind);
    mirorr::xf::Xf x = mirorr::xf::identity(kind);
    mirorr::xf::set_center(x, src.c);
    if (!mirorr::xf::set_matrix(x, src.M))
      throw IoError(MIRORR_ERR_ALGO, "Attempting to set a non-orthogonal rotation matrix "
                                      "(rigid transform from " + std::string(path) + ")");
    mirorr::xf::set_translation(x, src.t);
    mirorr::xf::get_params(x, params);
    for (int i = 0; i < 3; ++i) center[i] = x.c[i];
  });
}

}  // extern "C"
