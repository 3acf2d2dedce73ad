// mirorr: the command-line drop-in for the reference's `mirorr` executable
// (mirorr.cxx:96-680, MirorrPyramidWrapper::Update in
// itkMirorrPyramidWrapper.txx:596-860), running the registration on the
// MI355X through the C ABI of libmirorr_gpu.so.
//
// Same option names, short forms, defaults, positional order (moving, fixed,
// last-tfm, start-tfm), output-transform naming, --fresh rule and output
// images. Options that select work this build does not implement
// (--blockmetric mi, --tfm-type quat) fail with a message instead of being
// ignored; --nthreads and --use-gpu-bm are accepted and have no effect.
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mirorr_gpu.h"
#include "../../include/mirorr_io.h"
#include "../csrc/xform.hpp"

namespace xf = mirorr::xf;

namespace {

// ---- options -----------------------------------------------------------------
struct Opt {
  const char* name;
  char short_name;  // 0: none
  bool takes_value;
  const char* def;  // default value (nullptr: unset)
  const char* help;
};

const Opt kOptions[] = {
    {"moving", 'm', true, nullptr, "Filename of the moving image"},
    {"fixed", 'f', true, nullptr, "Filename of the fixed image"},
    {"last-tfm", 'l', true, "", "File to write the final transformation to"},
    {"start-tfm", 's', true, "", "File with initial transformation"},
    {"tfm-type", 't', true, "rigid", "Transformation type: rigid or affine"},
    {"reg-mode", 0, true, "symmetric", "Registration mode: symmetric or classic"},
    {"do-not-register", 'R', false, nullptr, "Do not run the registration"},
    {"save-reoriented-fixed", 0, true, nullptr, "Save the fixed image in moving space (header only if rigid)"},
    {"save-reoriented-moving", 0, true, nullptr, "Save the moving image in fixed space (header only if rigid)"},
    {"save-resampled-fixed", 0, true, nullptr, "Save the fixed image resampled in moving space"},
    {"save-resampled-moving", 0, true, nullptr, "Save the moving image resampled in fixed space"},
    {"final-interpolator", 0, true, "bspline", "Final interpolator: nn, linear, bspline or sinc"},
    {"moving-mask", 'M', true, "", "Mask of the moving image"},
    {"fixed-mask", 'F', true, "", "Mask of the fixed image"},
    {"switch-images", 'S', false, nullptr, "Switch the moving and fixed images"},
    {"invert-last-tfm", 'I', false, nullptr, "Invert the output transform"},
    {"invert-start-tfm", 'i', false, nullptr, "Invert the input transform"},
    {"nthreads", 0, true, "0", "Accepted for compatibility (the block matching runs on the GPU)"},
    {"echo-cmd", 0, false, nullptr, "Echo the command line"},
    {"fresh", 0, false, nullptr, "Do not initialise from an existing output transform"},
    {"help", 'h', false, nullptr, "Produce help message"},
    {"verbose", 'v', false, nullptr, "Request extra verbose output"},
    {"quiet", 'q', false, nullptr, "Request quiet output"},
    {"pyr-start", 'a', true, "1", "First pyramid level to process (-ve counts back from max)"},
    {"pyr-end", 'c', true, "1", "Last pyramid level to process (-ve counts back from max)"},
    {"pyr-num", 'd', true, "1024", "Maximum number of pyramid levels"},
    {"pyr-min-size", 'e', true, "32", "Minimum image size at the coarsest level"},
    {"resampling-mode", 0, true, "middle", "Resampling mode: middle, basic, moving or fixed"},
    {"crop", 0, true, "", "Crop the moving and fixed images on both sides of axes 0 and 1: \"mx,my,fx,fy\" voxels"},
    {"iterations", 'n', true, "0", "Iterations per level (0: default, -ve halves per level)"},
    {"portion-kept", 0, true, "0.5", "Portion of the block matches kept"},
    {"use-gpu-bm", 0, false, nullptr, "Accepted for compatibility (always on)"},
    {"reorient-fixed", 0, false, nullptr, "Shift the fixed image to identity space (ARI, origin 0) before registering"},
    {"reorient-moving", 0, false, nullptr, "Shift the moving image to identity space (ARI, origin 0) before registering"},
    {"save-fixed", 0, true, nullptr, "Save the fixed image resampled in moving space"},
    {"save-moving", 0, true, nullptr, "Save the moving image resampled in fixed space"},
    {"no-centering-init", '0', false, nullptr, "Do not centre the images before registration"},
    {"blockmetric", 0, true, "nc", "Block metric: nc, sd or cr"},
    {"npw-bins", 0, true, "32", "Accepted for compatibility"},
    {"npw-scale-factor", 0, true, "2", "Accepted for compatibility"},
    {"npw-shape-expansion", 0, false, nullptr, "Accepted for compatibility"},
    {"nhoodwidth", 0, true, "7", "Search neighbourhood width (candidates per axis)"},
    {"nhoodgap", 0, true, "3", "Gap between blocks"},
    {"blockwidth", 0, true, "4", "Block width in voxels"},
    {"blockgap", 0, true, "1", "Gap between candidate positions"},
};
const char* kPositional[] = {"moving", "fixed", "last-tfm", "start-tfm"};

struct Args {
  std::map<std::string, std::string> v;  // given or defaulted
  std::map<std::string, bool> given;
  bool has(const std::string& k) const { return given.count(k) > 0; }
  const std::string& str(const std::string& k) const {
    static const std::string empty;
    auto it = v.find(k);
    return it == v.end() ? empty : it->second;
  }
  int integer(const std::string& k) const {
    char* end = nullptr;
    const long x = std::strtol(str(k).c_str(), &end, 10);
    if (str(k).empty() || *end) throw std::runtime_error("option --" + k + " needs an integer");
    return (int)x;
  }
  double real(const std::string& k) const {
    char* end = nullptr;
    const double x = std::strtod(str(k).c_str(), &end);
    if (str(k).empty() || *end) throw std::runtime_error("option --" + k + " needs a number");
    return x;
  }
};

const Opt* find_long(const std::string& name) {
  const Opt* hit = nullptr;
  int n = 0;
  for (const Opt& o : kOptions) {
    if (name == o.name) return &o;
    if (std::strncmp(o.name, name.c_str(), name.size()) == 0) {
      hit = &o;
      ++n;
    }
  }
  if (n > 1) throw std::runtime_error("option '--" + name + "' is ambiguous");
  return hit;  // unique prefix, as program_options guesses
}

const Opt* find_short(char c) {
  for (const Opt& o : kOptions)
    if (o.short_name == c) return &o;
  return nullptr;
}

Args parse(int argc, char** argv) {
  Args a;
  for (const Opt& o : kOptions)
    if (o.def) a.v[o.name] = o.def;
  int pos = 0;
  auto set = [&](const Opt* o, const std::string& val) {
    if (a.has(o->name)) throw std::runtime_error(std::string("option '--") + o->name +
                                                 "' cannot be specified more than once");
    a.v[o->name] = val;
    a.given[o->name] = true;
  };
  for (int i = 1; i < argc; ++i) {
    std::string t = argv[i];
    if (t.size() > 2 && t[0] == '-' && t[1] == '-') {
      std::string name = t.substr(2), val;
      const size_t eq = name.find('=');
      const bool has_eq = eq != std::string::npos;
      if (has_eq) {
        val = name.substr(eq + 1);
        name = name.substr(0, eq);
      }
      const Opt* o = find_long(name);
      if (!o) throw std::runtime_error("unrecognised option '" + t + "'");
      if (o->takes_value && !has_eq) {
        if (i + 1 >= argc) throw std::runtime_error("option '--" + name + "' needs a value");
        val = argv[++i];
      }
      if (!o->takes_value && has_eq) throw std::runtime_error("option '--" + name + "' takes no value");
      set(o, val);
    } else if (t.size() > 1 && t[0] == '-' && !(std::isdigit((unsigned char)t[1]) && t.size() > 2) &&
               t[1] != '-') {
      const Opt* o = find_short(t[1]);
      if (!o) throw std::runtime_error("unrecognised option '" + t + "'");
      std::string val;
      if (o->takes_value) {
        if (t.size() > 2) {
          val = t.substr(2);
        } else {
          if (i + 1 >= argc) throw std::runtime_error(std::string("option '-") + t[1] + "' needs a value");
          val = argv[++i];
        }
      } else if (t.size() > 2) {
        throw std::runtime_error("unrecognised option '" + t + "'");
      }
      set(o, val);
    } else {
      if (pos >= 4) throw std::runtime_error("too many positional arguments ('" + t + "')");
      const Opt* o = find_long(kPositional[pos++]);
      set(o, t);
    }
  }
  return a;
}

void usage() {
  std::printf(
      "mirorr\n\n"
      "Synopsis: registers a 'moving' image to a 'fixed' image by block matching and robust\n"
      "regression (NCC blocks, least trimmed squares), on the GPU.\n"
      "Usage:\n  mirorr movingImage fixedImage [final.tfm] [start.tfm] ...\n\nOptions:\n");
  for (const Opt& o : kOptions) {
    char head[96];
    if (o.short_name)
      std::snprintf(head, sizeof(head), "  -%c [ --%s ]%s", o.short_name, o.name,
                    o.takes_value ? " arg" : "");
    else
      std::snprintf(head, sizeof(head), "  --%s%s", o.name, o.takes_value ? " arg" : "");
    std::printf("%-38s %s", head, o.help);
    if (o.def && *o.def) std::printf(" (=%s)", o.def);
    std::printf("\n");
  }
}

// ---- helpers -----------------------------------------------------------------
bool file_exists(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0;
}

// boost::filesystem basename(basename(leaf(p))): the file name without its
// last two extensions.
std::string stem2(const std::string& p) {
  std::string s = p.substr(p.find_last_of('/') == std::string::npos ? 0 : p.find_last_of('/') + 1);
  for (int k = 0; k < 2; ++k) {
    const size_t dot = s.find_last_of('.');
    if (dot != std::string::npos && dot > 0) s = s.substr(0, dot);
  }
  return s;
}

void check_io(mirorr_status st) {
  if (st != MIRORR_OK) throw std::runtime_error(mirorr_io_last_error());
}

struct Image {
  mirorr_geom g{};
  std::vector<float> data;
  std::vector<uint8_t> mask;
};

Image read_image(const std::string& path) {
  Image im;
  float* buf = nullptr;
  if (mirorr_nifti_read(path.c_str(), &im.g, &buf) != MIRORR_OK)
    throw std::runtime_error("cannot read image '" + path + "': " + mirorr_io_last_error());
  const size_t n = (size_t)im.g.dims[0] * im.g.dims[1] * im.g.dims[2];
  im.data.assign(buf, buf + n);
  mirorr_io_free(buf);
  return im;
}

void read_mask_into(Image& im, const std::string& path) {
  mirorr_geom g{};
  uint8_t* buf = nullptr;
  if (mirorr_nifti_read_mask(path.c_str(), &g, &buf) != MIRORR_OK)
    throw std::runtime_error("cannot read mask '" + path + "': " + mirorr_io_last_error());
  const size_t n = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
  for (int i = 0; i < 3; ++i)
    if (g.dims[i] != im.g.dims[i]) {
      mirorr_io_free(buf);
      throw std::runtime_error("mask '" + path + "' does not have the size of its image");
    }
  im.mask.assign(buf, buf + n);
  mirorr_io_free(buf);
}

std::string fmt_params(const double* p, int n) {
  std::string s = "[";
  char b[32];
  for (int i = 0; i < n; ++i) {
    std::snprintf(b, sizeof(b), "%s%.7f", i ? ", " : "", p[i]);
    s += b;
  }
  return s + "]";
}

// Rotation vector (axis * angle) of a direction matrix, for the warning on
// differing image directions (itkMirorrPyramidWrapper.txx:140-170, 300-309).
void rotation_vector(const double* d, double v[3]) {
  const double c = std::fmax(-1.0, std::fmin(1.0, (d[0] + d[4] + d[8] - 1.0) / 2.0));
  const double ang = std::acos(c);
  const double s = std::sin(ang);
  const double a[3] = {d[7] - d[5], d[2] - d[6], d[3] - d[1]};
  for (int i = 0; i < 3; ++i) v[i] = std::fabs(s) > 1e-12 ? a[i] * ang / (2.0 * s) : 0.0;
}

xf::Xf make_xf(bool affine, const double* params, const double* center) {
  xf::Xf x = xf::identity(affine ? xf::kAffine : xf::kRigid);
  xf::set_params(x, params);
  xf::set_center(x, center);
  return x;
}

// Header-only reorientation of a rigidly transformed image
// (itkMirorrPyramidImplement.txx:248-357 GetReorientedImage): direction
// columns mapped by the transform's matrix, the origin by the transform.
mirorr_geom reoriented_geom(const mirorr_geom& g, const xf::Xf& T) {
  mirorr_geom o = g;
  double D[3][3];
  for (int c = 0; c < 3; ++c) {
    double col[3], mc[3], n = 0.0;
    for (int r = 0; r < 3; ++r) col[r] = g.direction[3 * r + c] * g.spacing[c];
    xf::mulv(T.M, col, mc);
    for (int r = 0; r < 3; ++r) n += mc[r] * mc[r];
    n = std::sqrt(n);
    for (int r = 0; r < 3; ++r) D[r][c] = mc[r] / n;
  }
  const double det = std::fabs(D[0][0] * (D[1][1] * D[2][2] - D[1][2] * D[2][1]) -
                               D[0][1] * (D[1][0] * D[2][2] - D[1][2] * D[2][0]) +
                               D[0][2] * (D[1][0] * D[2][1] - D[1][1] * D[2][0]));
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) o.direction[3 * r + c] = D[r][c] / det;
  xf::apply(T, g.origin, o.origin);
  return o;
}

// GetTransformFromImageToIdentitySpace (itkMirorrUtilities.cxx:44-160,
// with do_centre off as the wrapper calls it): each row of the direction
// matrix snapped to its largest entry (+-1), divided by |det|, and the
// affine map of four grid landmarks (index 0 and the unit steps) from the
// image onto the same grid with that direction and origin 0. With the
// landmarks exact the map is M = D' D^-1, offset = -M o.
xf::Xf identity_space_tfm(const mirorr_geom& g) {
  xf::M3 D, Dp;
  for (int r = 0; r < 3; ++r) {
    int pos = 0;
    for (int c = 1; c < 3; ++c)
      if (std::fabs(g.direction[3 * r + c]) > std::fabs(g.direction[3 * r + pos])) pos = c;
    for (int c = 0; c < 3; ++c) {
      D.a[r][c] = g.direction[3 * r + c];
      Dp.a[r][c] = c == pos ? (g.direction[3 * r + c] > 0.0 ? 1.0 : -1.0) : 0.0;
    }
  }
  const double det = std::fabs(xf::det(Dp));
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Dp.a[r][c] /= det;
  xf::Xf x = xf::identity(xf::kAffine);
  x.M = xf::mul(Dp, xf::inverse(D));
  double mo[3];
  xf::mulv(x.M, g.origin, mo);
  for (int i = 0; i < 3; ++i) x.off[i] = -mo[i];
  xf::compute_translation(x);
  return x;
}

// ShiftImageToIdSpaceInplace (itkMirorrPyramidWrapper.txx:94-118): the ARI
// direction and origin 0.
void shift_to_identity_space(mirorr_geom& g) {
  const double ari[9] = {0, 1, 0, 1, 0, 0, 0, 0, 1};
  for (int i = 0; i < 9; ++i) g.direction[i] = ari[i];
  for (int i = 0; i < 3; ++i) g.origin[i] = 0.0;
}

// --crop "mx,my,fx,fy": voxels removed at both ends of axes 0 and 1 of the
// moving and fixed images (mirorr.cxx:205-207,414-420; parseInputToList
// splits on any of "x :,()=" and missing entries are 0).
std::vector<unsigned> parse_crop(const std::string& s) {
  std::vector<unsigned> out(4, 0u);
  std::vector<std::string> parts(1);
  for (char ch : s) {
    if (std::strchr("x :,()=", ch)) parts.emplace_back();
    else parts.back() += ch;
  }
  for (size_t i = 0; i < parts.size() && i < 4; ++i) {
    if (parts[i].empty()) continue;
    char* end = nullptr;
    const long v = std::strtol(parts[i].c_str(), &end, 10);
    if (*end || v < 0) throw std::runtime_error("Failed to parse input string: \"" + s + "\"");
    out[i] = (unsigned)v;
  }
  return out;
}

// CropImageFilter with lower = upper = (cx, cy, 0): the kept voxels keep
// their physical positions.
void crop_xy(Image& im, unsigned cx, unsigned cy) {
  const mirorr_geom g = im.g;
  if (2 * (int64_t)cx >= g.dims[0] || 2 * (int64_t)cy >= g.dims[1])
    throw std::runtime_error("--crop removes the whole image");
  mirorr_geom o = g;
  o.dims[0] = g.dims[0] - 2 * (int)cx;
  o.dims[1] = g.dims[1] - 2 * (int)cy;
  for (int r = 0; r < 3; ++r)
    o.origin[r] = g.origin[r] + g.direction[3 * r + 0] * g.spacing[0] * cx +
                  g.direction[3 * r + 1] * g.spacing[1] * cy;
  const size_t n = (size_t)o.dims[0] * o.dims[1] * o.dims[2];
  std::vector<float> d(n);
  std::vector<uint8_t> m(im.mask.empty() ? 0 : n);
  for (int z = 0; z < o.dims[2]; ++z)
    for (int y = 0; y < o.dims[1]; ++y)
      for (int x = 0; x < o.dims[0]; ++x) {
        const size_t src = (size_t)(x + cx) + (size_t)g.dims[0] * ((size_t)(y + cy) + (size_t)g.dims[1] * z);
        const size_t dst = (size_t)x + (size_t)o.dims[0] * ((size_t)y + (size_t)o.dims[1] * z);
        d[dst] = im.data[src];
        if (!m.empty()) m[dst] = im.mask[src];
      }
  im.g = o;
  im.data.swap(d);
  im.mask.swap(m);
}

int interpolator_code(const std::string& name) {
  if (name.find("nn") != std::string::npos) return MIRORR_INTERP_NN;
  if (name.find("linear") != std::string::npos) return MIRORR_INTERP_LINEAR;
  if (name.find("bspline") != std::string::npos) return MIRORR_INTERP_BSPLINE;
  if (name.find("sinc") != std::string::npos) return MIRORR_INTERP_SINC;
  std::printf("GetInterpolatorFromString(...) Unspecified or unknown interpolator name. "
              "Defaulting to bspline.\n");
  return MIRORR_INTERP_BSPLINE;
}

struct Ctx {
  mirorr_ctx* c = nullptr;
  ~Ctx() {
    if (c) mirorr_ctx_destroy(c);
  }
  void check(mirorr_status st) const {
    if (st != MIRORR_OK) throw std::runtime_error(mirorr_last_error(c));
  }
};

// Resampled output (GetResampledImage): `image` on the grid of `ref`,
// out(p) = image(T p).
void write_resampled(Ctx& ctx, const std::string& path, const Image& image, const Image& ref,
                     const xf::Xf& T, int interp) {
  if (!ctx.c && mirorr_ctx_create(0, 0, &ctx.c) != MIRORR_OK)
    throw std::runtime_error("cannot create a GPU context");
  std::vector<float> out((size_t)ref.g.dims[0] * ref.g.dims[1] * ref.g.dims[2]);
  double M[9];
  for (int i = 0; i < 9; ++i) M[i] = T.M.a[i / 3][i % 3];
  ctx.check(mirorr_resample_final(ctx.c, image.data.data(), &image.g, M, T.off, &ref.g, interp,
                                  out.data()));
  check_io(mirorr_nifti_write(path.c_str(), &ref.g, out.data()));
}

int run(int argc, char** argv) {
  const Args a = parse(argc, argv);
  const time_t t0 = std::time(nullptr);
  std::printf(
      "------------------------------------------------------------------------\n"
      "- mirorr (MI355X block-matching registration)                         -\n"
      "------------------------------------------------------------------------\n"
      "Current time is:  %s", std::asctime(std::localtime(&t0)));
  if (a.has("echo-cmd")) {
    for (int i = 0; i < argc; ++i) std::printf("%s ", argv[i]);
    std::printf("\n");
  } else {
    std::printf("(use --echo-cmd to get more process information)\n");
  }
  std::printf("------------------------------------------------------------------------\n\n");
  if (a.has("help")) {
    usage();
    return 1;
  }
  if (!a.has("moving")) {
    std::fprintf(stderr, "No moving image supplied. Type 'mirorr --help' for usage.\n");
    return -1;
  }
  if (!a.has("fixed")) {
    std::fprintf(stderr, "No fixed image supplied. Type 'mirorr --help' for usage.\n");
    return -1;
  }
  for (const char* k : {"moving", "fixed"})
    if (!file_exists(a.str(k))) {
      std::fprintf(stderr, "ERROR: %s not found. Exiting!\n", a.str(k).c_str());
      return -1;
    }
  const std::string mode = a.str("reg-mode");
  if (mode != "classic" && mode != "symmetric") {
    std::fprintf(stderr, "ERROR: Unknown reg-mode. Available options are 'classic' and "
                         "'symmetric' (default).\n");
    return -1;
  }
  const bool reorient_fixed = a.has("reorient-fixed"), reorient_moving = a.has("reorient-moving");
  const std::vector<unsigned> crop = parse_crop(a.str("crop"));
  // mirorr.cxx:532-555. max-resolution and max-size pass the option parser
  // but StartOptimization rejects them (itkMirorrRegistrationMethod.txx:573-579).
  const std::string rmode = a.str("resampling-mode");
  int resampling = MIRORR_RESAMPLE_MIDDLE;
  if (rmode == "basic") {
    resampling = MIRORR_RESAMPLE_BASIC;
  } else if (rmode == "moving") {
    resampling = MIRORR_RESAMPLE_MOVING;
  } else if (rmode == "fixed") {
    resampling = MIRORR_RESAMPLE_FIXED;
  } else if (rmode == "max-resolution" || rmode == "max-size") {
    throw std::runtime_error("Unknown resampling type (" + rmode + "). This should not happen...");
  } else if (rmode != "middle") {
    std::fprintf(stderr, "ERROR: Unknown resampling-mode. Available options are 'classic' and "
                         "'symmetric' (default).\n");
    return -1;
  }
  const std::string metric = a.str("blockmetric");
  int block_metric = MIRORR_METRIC_NC;
  if (metric == "sd") {
    block_metric = MIRORR_METRIC_SSD;
  } else if (metric == "cr") {
    block_metric = MIRORR_METRIC_CR;
  } else if (metric == "mi") {
    throw std::runtime_error("--blockmetric mi is not supported by this build (nc, sd or cr)");
  } else if (metric != "nc") {
    std::fprintf(stderr, "ERROR: %s is not a valid blockmetric type. Exiting!\n"
                         "possible options are:\n  cr: correlation_ratio\n"
                         "  mi: mutual_information\n  nc: normalized_correlation\n"
                         "  sd: sum_of_squared_differences\n", metric.c_str());
    return -1;
  }
  const std::string tfm_type = a.str("tfm-type");
  if (tfm_type != "rigid" && tfm_type != "affine")
    throw std::runtime_error("--tfm-type " + tfm_type + " is not supported by this build "
                             "(rigid or affine)");
  const bool affine = tfm_type == "affine";
  const int np = affine ? 12 : 6;

  const bool sw = a.has("switch-images");
  const std::string moving_name = sw ? a.str("fixed") : a.str("moving");
  const std::string fixed_name = sw ? a.str("moving") : a.str("fixed");
  const std::string moving_mask = sw ? a.str("fixed-mask") : a.str("moving-mask");
  const std::string fixed_mask = sw ? a.str("moving-mask") : a.str("fixed-mask");
  const bool do_not_register = a.has("do-not-register");

  std::string out_tfm = a.str("last-tfm"), in_tfm = a.str("start-tfm");
  if (out_tfm.empty() && !do_not_register) {
    out_tfm = stem2(fixed_name) + "_tfmFinal.tfm";
    std::fprintf(stderr, "WARNING: No output transform supplied. Setting to: %s\n",
                 out_tfm.c_str());
    std::fprintf(stderr, "INFO:    To avoid saving the last transform, use '--last-tfm ignore'\n");
  } else if (out_tfm == "ignore") {
    out_tfm.clear();
  }
  if (in_tfm.empty() && !out_tfm.empty() && file_exists(out_tfm)) {
    std::printf("INFO: Output transform file already exists (%s) and will be overwritten, ",
                out_tfm.c_str());
    if (a.has("fresh")) {
      std::printf("and it will NOT be used to initialise registration.\n\n");
    } else {
      std::printf("but it WILL be used to initialise registration.\n\n");
      in_tfm = out_tfm;
    }
  }
  std::string save_fixed, save_moving;
  bool force_fixed = false, force_moving = false;
  if (a.has("save-fixed")) save_fixed = a.str("save-fixed"), force_fixed = true;
  if (a.has("save-moving")) save_moving = a.str("save-moving"), force_moving = true;
  if (a.has("save-reoriented-fixed")) save_fixed = a.str("save-reoriented-fixed"), force_fixed = false;
  if (a.has("save-reoriented-moving")) save_moving = a.str("save-reoriented-moving"), force_moving = false;
  if (a.has("save-resampled-fixed")) {
    if (!save_fixed.empty())
      throw std::runtime_error("Cannot use save-resampled-fixed and save-reoriented-fixed arguments at same time.");
    save_fixed = a.str("save-resampled-fixed");
    force_fixed = true;
  }
  if (a.has("save-resampled-moving")) {
    if (!save_moving.empty())
      throw std::runtime_error("Cannot use save-resampled-moving and save-reoriented-moving arguments at same time.");
    save_moving = a.str("save-resampled-moving");
    force_moving = true;
  }
  const int interp = interpolator_code(a.str("final-interpolator"));
  int verbosity = 1;
  if (a.has("quiet")) verbosity = 0;
  if (a.has("verbose")) verbosity = 2;

  mirorr_reg_params rp;
  mirorr_reg_default_params(&rp);
  rp.symmetric = mode == "symmetric";
  rp.affine = affine;
  rp.iterations = a.integer("iterations");
  rp.pyr_start = a.integer("pyr-start");
  rp.pyr_end = a.integer("pyr-end");
  rp.pyr_num = a.integer("pyr-num");
  rp.pyr_min_size = a.integer("pyr-min-size");
  rp.portion_kept = a.real("portion-kept");
  rp.verbosity = verbosity;
  rp.centre_init = !a.has("no-centering-init");
  rp.block_metric = block_metric;
  rp.resampling = resampling;
  if (a.integer("blockwidth") > 0) rp.block_width = a.integer("blockwidth");
  if (a.integer("blockgap") > 0) rp.block_gap = a.integer("blockgap");
  if (a.integer("nhoodwidth") > 0) rp.nhood_width = a.integer("nhoodwidth");
  if (a.integer("nhoodgap") > 0) rp.nhood_gap = a.integer("nhoodgap");

  // ---- MirorrPyramidWrapper::Update -------------------------------------------
  Image moving = read_image(moving_name), fixed = read_image(fixed_name);
  const Image moving_in = moving, fixed_in = fixed;  // as read, for the outputs
  if (!moving_mask.empty()) read_mask_into(moving, moving_mask);
  if (!fixed_mask.empty()) read_mask_into(fixed, fixed_mask);
  check_io(mirorr_reorient_ari(&fixed.g, fixed.data.data(), fixed.mask.empty() ? nullptr : fixed.mask.data()));
  check_io(mirorr_reorient_ari(&moving.g, moving.data.data(), moving.mask.empty() ? nullptr : moving.mask.data()));
  // Transforms to identity space, taken before any crop
  // (itkMirorrPyramidWrapper.txx:213-240), then the optional shifts and crops.
  const xf::Xf fixed_to_id = identity_space_tfm(fixed.g);
  const xf::Xf moving_to_id = identity_space_tfm(moving.g);
  if (reorient_fixed) {
    std::printf("Fixed: %s\n", fixed_name.c_str());
    shift_to_identity_space(fixed.g);
  }
  if (reorient_moving) {
    std::printf("Moving: %s\n", moving_name.c_str());
    shift_to_identity_space(moving.g);
  }
  if (crop[0] > 0 || crop[1] > 0) crop_xy(moving, crop[0], crop[1]);
  if (crop[2] > 0 || crop[3] > 0) crop_xy(fixed, crop[2], crop[3]);
  {
    double va[3], vb[3];
    rotation_vector(moving.g.direction, va);
    rotation_vector(fixed.g.direction, vb);
    const double delta = std::sqrt((va[0] - vb[0]) * (va[0] - vb[0]) +
                                   (va[1] - vb[1]) * (va[1] - vb[1]) +
                                   (va[2] - vb[2]) * (va[2] - vb[2]));
    if (delta > 0.017453)
      std::fprintf(stderr,
                   "#\n# Warning: The moving and fixed images have different cosine direction "
                   "matrices.\n#          Low quality results are to be expected.\n"
                   "#          Note: DELTA = %g.\n#\n", delta);
  }

  double params[12] = {0}, center[3] = {0, 0, 0};
  const int kind = affine ? xf::kAffine : xf::kRigid;
  xf::Xf init = xf::identity(kind);
  xf::get_params(init, params);
  const bool did_init = in_tfm.empty() && rp.centre_init;
  if (did_init) {
    check_io(mirorr_centered_init(&moving.g, &fixed.g, affine, params, center));
  } else if (!in_tfm.empty()) {
    const mirorr_status st = mirorr_tfm_read(in_tfm.c_str(), affine, params, center,
                                             a.has("invert-start-tfm"));
    if (st == MIRORR_ERR_ARG && std::strstr(mirorr_io_last_error(), "MatrixOffsetTransform"))
      throw std::runtime_error(std::string("Error: ") + mirorr_io_last_error());
    if (st != MIRORR_OK) {
      const std::string example = std::string("__tmp_example-") +
                                  (affine ? "AffineTransform" : "Euler3DTransform") + "_double_3_3.tfm";
      std::fprintf(stderr,
                   "Failed to read %s using as ITK transform file, but failed, because %s\n"
                   "Continuing.\nCould not read input transform file: %s\n"
                   "An example transform file has been written here:\n\n%s\n\n"
                   "Try editing this file and using it as an input.\n",
                   in_tfm.c_str(), mirorr_io_last_error(), in_tfm.c_str(), example.c_str());
      xf::get_params(init, params);
      for (double& c : center) c = 0.0;
      check_io(mirorr_tfm_write(example.c_str(), affine, params, center, 0));
    }
  }
  // The start transform, composed into an identity transform with centre 0
  // between the identity-space shifts of the reoriented images
  // (itkMirorrPyramidWrapper.txx:626-648).
  {
    const xf::Xf orig = make_xf(affine, params, center);
    xf::Xf t = xf::identity(kind);
    if (reorient_moving && !did_init) xf::compose(t, xf::inverse(moving_to_id), false);
    xf::compose(t, orig, false);
    if (reorient_fixed && !did_init) xf::compose(t, fixed_to_id, false);
    xf::get_params(t, params);
    for (int i = 0; i < 3; ++i) center[i] = t.c[i];
  }
  if (verbosity >= 1)
    std::printf("Transform Class:                 %s\n"
                "Transformation parameters:       %s\n"
                "Transformation fixed parameters: %s (usually this is the rotation center)\n",
                affine ? "AffineTransform" : "Euler3DTransform", fmt_params(params, np).c_str(),
                fmt_params(center, 3).c_str());

  Ctx ctx;
  double out_params[12] = {0}, out_center[3] = {0, 0, 0};
  if (!do_not_register) {
    if (mirorr_ctx_create(0, 0, &ctx.c) != MIRORR_OK)
      throw std::runtime_error("cannot create a GPU context");
    mirorr_volume *vf = nullptr, *vm = nullptr;
    ctx.check(mirorr_volume_create(ctx.c, fixed.data.data(), fixed.mask.empty() ? nullptr : fixed.mask.data(), &fixed.g, &vf));
    const mirorr_status sm = mirorr_volume_create(ctx.c, moving.data.data(), moving.mask.empty() ? nullptr : moving.mask.data(), &moving.g, &vm);
    if (sm != MIRORR_OK) {
      mirorr_volume_destroy(vf);
      ctx.check(sm);
    }
    mirorr_reg_stats stats{};
    const mirorr_status sr = mirorr_register(ctx.c, &rp, vf, vm, params, center, out_params,
                                             out_center, &stats);
    mirorr_volume_destroy(vm);
    mirorr_volume_destroy(vf);
    ctx.check(sr);
    if (verbosity >= 1)
      std::printf("Registration: %d levels, %.3f s on the GPU\n", stats.levels, stats.seconds);
  } else {
    std::printf("\nINFO: Not running registration as requested.\n");
    std::memcpy(out_params, params, sizeof(params));
    std::memcpy(out_center, center, sizeof(center));
  }

  // The written transform: the identity-space shifts and the registration
  // transform composed into an identity transform with centre 0
  // (itkMirorrPyramidWrapper.txx:680-700).
  xf::Xf T = xf::identity(kind);
  if (reorient_moving) xf::compose(T, moving_to_id, false);
  xf::compose(T, make_xf(affine, out_params, out_center), false);
  if (reorient_fixed) xf::compose(T, xf::inverse(fixed_to_id), false);
  double full[12];
  xf::get_params(T, full);
  const double zero[3] = {0, 0, 0};
  if (!out_tfm.empty()) check_io(mirorr_tfm_write(out_tfm.c_str(), affine, full, zero, a.has("invert-last-tfm")));
  std::printf("   RegTfm: %s %s\n", fmt_params(out_params, np).c_str(), fmt_params(out_center, 3).c_str());
  std::printf("  FullTfm: %s %s\n", fmt_params(full, np).c_str(), fmt_params(zero, 3).c_str());

  const bool rigid = !affine;
  const xf::Xf Tinv = xf::inverse(T);
  if (!save_fixed.empty()) {
    if (!force_fixed && rigid) {
      const mirorr_geom g = reoriented_geom(fixed_in.g, Tinv);
      check_io(mirorr_nifti_write(save_fixed.c_str(), &g, fixed_in.data.data()));
    } else {
      write_resampled(ctx, save_fixed, fixed_in, moving_in, T, interp);
    }
  }
  if (!save_moving.empty()) {
    if (!force_moving && rigid) {
      const mirorr_geom g = reoriented_geom(moving_in.g, T);
      check_io(mirorr_nifti_write(save_moving.c_str(), &g, moving_in.data.data()));
    } else {
      write_resampled(ctx, save_moving, moving_in, fixed_in, Tinv, interp);
    }
  }
  const time_t t1 = std::time(nullptr);
  std::printf("\nmirorr ran successfully\nProcess started on:   %s", std::asctime(std::localtime(&t0)));
  std::printf("Process completed on: %s", std::asctime(std::localtime(&t1)));
  std::printf("Elapsed time is %lds\n", (long)(t1 - t0));
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fflush(stdout);
    std::fprintf(stderr, "ERROR: %s\n", e.what());
    return 1;
  }
}
