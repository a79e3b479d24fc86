// Host engine (see srs_engine.h).  Arithmetic that the GPU integrals depend on
// bit for bit (field, RK4, particle set-up, grids) keeps the reference's
// operation order; this file is compiled without FMA contraction and without
// -march so that it rounds exactly as the reference's x86-64 build does.
#include "srs_engine.h"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>

#include "srs_hip.h"

namespace srs {
namespace {

// Reference constants, TOSCARSSR.h:20-69.
const double kPi = 3.14159265358979323846;
const double kTwoPi = 2.0 * kPi;
const double kC = 2.99792458e8;
const double kQe = 1.602176462e-19;
const double kMe = 9.10938356e-31;

inline double kg_to_gev(double m) { return 1e-9 * m * kC * kC / kQe; }
inline double gev_to_kg(double m) { return kQe * m / (1e-9 * kC * kC); }

inline Vec3 sub(Vec3 const& a, Vec3 const& b) { return Vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 vadd(Vec3 const& a, Vec3 const& b) { return Vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline double mag2(Vec3 const& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double mag(Vec3 const& a) { return std::sqrt(mag2(a)); }
inline double dot(Vec3 const& a, Vec3 const& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 scale(Vec3 const& a, double s) { return Vec3(a.x * s, a.y * s, a.z * s); }
inline Vec3 divide(Vec3 const& a, double d) { return Vec3(a.x / d, a.y / d, a.z / d); }
inline Vec3 unit(Vec3 const& a) { return divide(a, mag(a)); }
inline Vec3 cross(Vec3 const& a, Vec3 const& b) {
  return Vec3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
inline void rotate_x(Vec3& v, double a) {
  double const s = std::sin(a), c = std::cos(a), yy = v.y;
  v.y = c * yy - s * v.z;
  v.z = s * yy + c * v.z;
}
inline void rotate_y(Vec3& v, double a) {
  double const s = std::sin(a), c = std::cos(a), zz = v.z;
  v.z = c * zz - s * v.x;
  v.x = s * zz + c * v.x;
}
inline void rotate_z(Vec3& v, double a) {
  double const s = std::sin(a), c = std::cos(a), xx = v.x;
  v.x = c * xx - s * v.y;
  v.y = s * xx + c * v.y;
}
inline void rotate_xyz(Vec3& v, Vec3 const& r) {
  rotate_x(v, r.x);
  rotate_y(v, r.y);
  rotate_z(v, r.z);
}
inline std::string lower(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), ::tolower);
  return s;
}
inline std::string upper(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), ::toupper);
  return s;
}

void check_status(int rc) {
  if (rc == SRS_OK) return;
  std::string msg = srs_hip_last_error();
  if (rc == SRS_ERR_INVALID) throw std::length_error(msg);
  throw std::invalid_argument(msg);
}

srs_traj view(Trajectory const& t) {
  srs_traj v;
  v.n = (int64_t)t.size();
  v.dt = t.dt;
  v.x = t.x.data();
  v.y = t.y.data();
  v.z = t.z.data();
  v.bx = t.bx.data();
  v.by = t.by.data();
  v.bz = t.bz.data();
  v.ax = t.ax.data();
  v.ay = t.ay.data();
  v.az = t.az.data();
  return v;
}

}  // namespace

// ---- field ---------------------------------------------------------------------

Vec3 IdealUndulator::eval(Vec3 const& X) const {
  if (kind != FieldKind::Undulator) {
    // TField3D_Gaussian::GetF (.cc:88-113), TField3D_UniformBox::GetF (.cc:94-117) and
    // TField3D_Grid::GetF (.cc:124-131): rotate the position, then subtract the centre
    // (the translation)
    Vec3 Xr = X;
    rotate_xyz(Xr, rot);
    Vec3 const RX = sub(Xr, center);
    if (kind == FieldKind::Grid) return grid->eval(RX);
    if (kind == FieldKind::Gaussian) {
      double frac = 1;
      if (width.x > 0) frac *= std::exp(-std::pow(RX.x / width.x, 2) / 2.);
      if (width.y > 0) frac *= std::exp(-std::pow(RX.y / width.y, 2) / 2.);
      if (width.z > 0) frac *= std::exp(-std::pow(RX.z / width.z, 2) / 2.);
      return Vec3(frac * field.x, frac * field.y, frac * field.z);
    }
    if ((width.x > 0 && std::fabs(RX.x) > std::fabs(width.x / 2.)) ||
        (width.y > 0 && std::fabs(RX.y) > std::fabs(width.y / 2.)) ||
        (width.z > 0 && std::fabs(RX.z) > std::fabs(width.z / 2.)))
      return Vec3(0, 0, 0);
    return field;
  }
  double const lam = mag(period);
  Vec3 const u = divide(period, mag(period));
  double const len = mag(period) * (nperiods + 2);
  if (!pole_delta.empty()) {
    // interior half-period poles scaled by (1 + delta_p); phase and taper unused
    double const D = dot(X, u);
    double const s = std::sin(2.0 * kPi * D / lam);
    if (D > len / 2. || D < -len / 2.) return Vec3();
    if (D < -len / 2. + lam || D > len / 2. - lam) {
      if (D < -len / 2. + lam / 2. || D > len / 2. - lam / 2.)
        return Vec3((0.25 * field.x) * s, (0.25 * field.y) * s, (0.25 * field.z) * s);
      return Vec3((0.75 * field.x) * s, (0.75 * field.y) * s, (0.75 * field.z) * s);
    }
    int p = (int)std::floor((D - (-len / 2. + lam)) / (lam / 2.));
    if (p < 0) p = 0;
    if (p > 2 * nperiods - 1) p = 2 * nperiods - 1;
    double const sc = 1.0 + pole_delta[(size_t)p];
    return Vec3((field.x * s) * sc, (field.y * s) * sc, (field.z * s) * sc);
  }
  // TField3D_IdealUndulator::GetF, .cc:106-139
  double const D = dot(sub(X, center), u);
  double const T = 1 + D * taper;
  double const PS = phase * mag(period) / kTwoPi;
  if (D > len / 2. + PS || D < -len / 2. + PS) return Vec3();
  double const s = std::sin(kTwoPi * (D - PS) / lam);
  if (D < -len / 2. + PS + lam || D > len / 2. + PS - lam) {
    double const f = (D < -len / 2. + PS + lam / 2. || D > len / 2. + PS - lam / 2.) ? 0.25 : 0.75;
    return Vec3(((field.x * f) * s) * T, ((field.y * f) * s) * T, ((field.z * f) * s) * T);
  }
  return Vec3((field.x * s) * T, (field.y * s) * T, (field.z * s) * T);
}

// ---- field maps (TField3D_Grid) ---------------------------------------------------

Vec3 FieldGrid::eval(Vec3 const& X) const {
  // TField3D_Grid::GetF, TField3D_Grid.cc:124-250: zero outside (open interval) on every
  // axis with more than one point; cell index by truncation, offset d = (x - start) - i*step;
  // each linear step is a + (d*(b - a))/step, x first, then y, then z.
  size_t i[3] = {0, 0, 0};
  double d[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a) {
    if (n[a] <= 1) continue;
    if (!(X[a] > start[a] && X[a] < stop[a])) return Vec3(0, 0, 0);  // (also for a NaN position)
    i[a] = (size_t)((X[a] - start[a]) / step[a]);
    // a point within rounding of the upper edge would index one past the grid (the
    // reference reads beyond its array there); stay in the last cell
    if (i[a] > (size_t)(n[a] - 2)) i[a] = (size_t)(n[a] - 2);
    d[a] = (X[a] - start[a]) - i[a] * step[a];
  }
  auto at = [&](size_t ix, size_t iy, size_t iz) {
    size_t const k = 3 * ((ix * (size_t)n[1] + iy) * (size_t)n[2] + iz);
    return Vec3(data[k], data[k + 1], data[k + 2]);
  };
  auto lerp = [](Vec3 const& a, Vec3 const& b, double dd, double st) {
    return Vec3(a.x + (dd * (b.x - a.x)) / st, a.y + (dd * (b.y - a.y)) / st, a.z + (dd * (b.z - a.z)) / st);
  };
  bool const hx = n[0] > 1, hy = n[1] > 1, hz = n[2] > 1;
  size_t const x0 = i[0], y0 = i[1], z0 = i[2];
  if (hx && hy && hz) {
    Vec3 const v00 = lerp(at(x0, y0, z0), at(x0 + 1, y0, z0), d[0], step[0]);
    Vec3 const v10 = lerp(at(x0, y0 + 1, z0), at(x0 + 1, y0 + 1, z0), d[0], step[0]);
    Vec3 const v01 = lerp(at(x0, y0, z0 + 1), at(x0 + 1, y0, z0 + 1), d[0], step[0]);
    Vec3 const v11 = lerp(at(x0, y0 + 1, z0 + 1), at(x0 + 1, y0 + 1, z0 + 1), d[0], step[0]);
    return lerp(lerp(v00, v10, d[1], step[1]), lerp(v01, v11, d[1], step[1]), d[2], step[2]);
  }
  if (hx && hy) return lerp(lerp(at(x0, y0, 0), at(x0 + 1, y0, 0), d[0], step[0]),
                            lerp(at(x0, y0 + 1, 0), at(x0 + 1, y0 + 1, 0), d[0], step[0]), d[1], step[1]);
  if (hx && hz) return lerp(lerp(at(x0, 0, z0), at(x0 + 1, 0, z0), d[0], step[0]),
                            lerp(at(x0, 0, z0 + 1), at(x0 + 1, 0, z0 + 1), d[0], step[0]), d[2], step[2]);
  if (hy && hz) return lerp(lerp(at(0, y0, z0), at(0, y0 + 1, z0), d[1], step[1]),
                            lerp(at(0, y0, z0 + 1), at(0, y0 + 1, z0 + 1), d[1], step[1]), d[2], step[2]);
  // one axis: the data are the points along it (index = the axis index)
  if (hx) return lerp(Vec3(data[3 * x0], data[3 * x0 + 1], data[3 * x0 + 2]),
                      Vec3(data[3 * x0 + 3], data[3 * x0 + 4], data[3 * x0 + 5]), d[0], step[0]);
  if (hy) return lerp(Vec3(data[3 * y0], data[3 * y0 + 1], data[3 * y0 + 2]),
                      Vec3(data[3 * y0 + 3], data[3 * y0 + 4], data[3 * y0 + 5]), d[1], step[1]);
  return lerp(Vec3(data[3 * z0], data[3 * z0 + 1], data[3 * z0 + 2]),
              Vec3(data[3 * z0 + 3], data[3 * z0 + 4], data[3 * z0 + 5]), d[2], step[2]);
}

namespace {

// First number on a header line (istream extraction, as TField3D_Grid::GetHeaderValue).
double header_value(std::string const& L) {
  std::istringstream S(L);
  double v = 0;
  if (!(S >> v)) throw std::invalid_argument("cannot read header value");
  return v;
}
// A header line that starts with the comment character, then the number (SRW format,
// TField3D_Grid::GetHeaderValueSRW).
double header_value_srw(std::string const& L, char comment) {
  std::istringstream S(L);
  char c = 0;
  S.get(c);
  if (c != comment) throw std::invalid_argument("something is wrong with the comment character, it was not seen");
  double v = 0;
  if (!(S >> v)) throw std::invalid_argument("cannot read header value SRW format");
  return v;
}
std::string next_line(std::istream& f) {
  std::string L;
  std::getline(f, L);
  return L;
}
void check_points(int nx, int ny, int nz) {
  // the reference tests NX, NY and NY (not NZ); an NZ < 1 grid has no data either
  if (nx < 1 || ny < 1 || nz < 1) throw std::out_of_range("invalid number of points in at least one dimension");
  if (nx == 1 && ny == 1 && nz == 1) throw std::out_of_range("invalid dimensions");
}
void finish_axes(FieldGrid& g) {
  for (int a = 0; a < 3; ++a) g.stop[a] = g.start[a] + (g.n[a] - 1) * g.step[a];
}
// Header of the OSCARS format scaled as TField3D_Grid::ReadFile does (.cc:318-400): the
// grid keeps its middle point while the steps scale.
void oscars_axes(FieldGrid& g, double const hv[10], std::vector<double> const& scale) {
  for (int a = 0; a < 3; ++a) {
    double const start_in = hv[1 + 3 * a], step_in = hv[2 + 3 * a];
    int const n = (int)hv[3 + 3 * a];
    double const step = scale.size() > (size_t)a ? step_in * scale[a] : step_in;
    double const middle = start_in + step_in * (n - 1) / 2.;
    g.n[a] = n;
    g.step[a] = step;
    g.start[a] = middle - step * (n - 1) / 2.;
  }
  check_points(g.n[0], g.n[1], g.n[2]);
  finish_axes(g);
}
void push_rotated(std::vector<double>& out, double fx, double fy, double fz, Vec3 const& rot) {
  Vec3 F(fx, fy, fz);
  rotate_xyz(F, rot);
  out.push_back(F.x);
  out.push_back(F.y);
  out.push_back(F.z);
}

FieldGrid read_oscars(std::string const& file, Vec3 const& rot, std::vector<double> const& scale) {
  std::ifstream f(file);
  if (!f) throw std::invalid_argument("cannot open file for reading");
  next_line(f);  // comment
  double hv[10] = {0};
  for (int k = 1; k < 10; ++k) hv[k] = header_value(next_line(f));
  FieldGrid g;
  oscars_axes(g, hv, scale);
  double const sx = scale.size() > 3 ? scale[3] : 1, sy = scale.size() > 4 ? scale[4] : 1,
               sz = scale.size() > 5 ? scale[5] : 1;
  size_t const np = (size_t)g.n[0] * g.n[1] * g.n[2];
  g.data.reserve(3 * np);
  std::string L;
  for (size_t k = 0; k < np; ++k) {
    std::getline(f, L);
    if (f.eof()) throw std::invalid_argument("error reading file.  Check format");
    std::istringstream S(L);
    double fx, fy, fz;
    S >> fx >> fy >> fz;
    if (sx != 1) fx *= sx;
    if (sy != 1) fy *= sy;
    if (sz != 1) fz *= sz;
    if (S.fail()) throw std::invalid_argument("error reading file.  Check format");
    push_rotated(g.data, fx, fy, fz, rot);
  }
  return g;
}

// OSCARS1D: columns named by the format string after its first 8 characters (one of
// X/Y/Z and any of Bx/By/Bz, Ex..., Fx...), sorted by position and resampled to 10000
// points per metre by linear interpolation (TField3D_Grid.cc:533-831).
FieldGrid read_oscars1d(std::string const& file, std::string const& format, Vec3 const& rot,
                        std::vector<double> const& scale, char comment) {
  std::ifstream f(file);
  if (!f) throw std::invalid_argument("cannot open file for reading");
  int order[4] = {-1, -1, -1, -1};
  std::istringstream fs(format.size() > 8 ? format.substr(8) : std::string());
  std::string c, axis;
  int index = 0, xdim = 0, fdim = 0;
  bool has[4] = {false, false, false, false};
  while (fs >> c) {
    if (index > 3) throw std::out_of_range("spatial or B-field dimensions are too large(index>3)");
    int slot;
    if (c == "X" || c == "Y" || c == "Z") {
      axis = c;
      ++xdim;
      slot = 0;
    } else if (c == "Bx" || c == "Ex" || c == "Fx") {
      slot = 1;
    } else if (c == "By" || c == "Ey" || c == "Fy") {
      slot = 2;
    } else if (c == "Bz" || c == "Ez" || c == "Fz") {
      slot = 3;
    } else {
      throw std::invalid_argument("only excepts X Y Z Bx By Bz");
    }
    if (slot > 0) {
      has[slot] = true;
      ++fdim;
    }
    order[index++] = slot;
  }
  if (xdim != 1) throw std::out_of_range("spatial or B-field dimensions are too large");
  int const ncols = xdim + fdim;
  std::vector<std::array<double, 4>> in;
  std::string L;
  std::getline(f, L);  // comment
  while (std::getline(f, L)) {
    size_t const first = L.find_first_not_of(" \t");
    if (first == std::string::npos || L[first] == comment) continue;
    std::istringstream S(L);
    double v[4] = {0, 0, 0, 0};
    for (int i = 0; i < ncols; ++i) {
      S >> v[order[i]];
      if (S.fail()) throw std::length_error("something is incorrect with data format or iformat string");
    }
    for (size_t k = 0; k < scale.size() && k < 4; ++k) v[order[k]] *= scale[k];
    in.push_back({{v[0], v[1], v[2], v[3]}});
  }
  std::sort(in.begin(), in.end(),
            [](std::array<double, 4> const& a, std::array<double, 4> const& b) { return a[0] < b[0]; });
  if (in.size() < 2) throw std::length_error("not enough data points");
  double const first = in.front()[0], last = in.back()[0];
  size_t const np = (size_t)((last - first) * 10000);
  double const step = (last - first) / (double)(np - 1);
  FieldGrid g;
  int const a = axis == "X" ? 0 : (axis == "Y" ? 1 : 2);
  g.n[a] = (int)np;
  g.start[a] = first;
  g.stop[a] = last;
  g.step[a] = step;
  if (np < 2) throw std::out_of_range("invalid dimensions");
  g.data.reserve(3 * np);
  size_t minbin = 0, after = 0, before = 0;
  for (size_t i = 0; i != np; ++i) {
    double const z = i * step + first;
    for (size_t j = minbin + 1; j != in.size(); ++j)
      if (in[j][0] > z) {
        after = j;
        before = j - 1;
        minbin = j - 1;
        break;
      }
    double const dz = in[after][0] - in[before][0];
    double const sx = !has[1] ? 0 : (in[after][1] - in[before][1]) / dz;
    double const sy = !has[2] ? 0 : (in[after][2] - in[before][2]) / dz;
    double const sz = !has[3] ? 0 : (in[after][3] - in[before][3]) / dz;
    double const bx = !has[1] ? 0 : in[before][1] + (z - in[before][0]) * sx;
    double const by = !has[2] ? 0 : in[before][2] + (z - in[before][0]) * sy;
    double const bz = !has[3] ? 0 : in[before][3] + (z - in[before][0]) * sz;
    push_rotated(g.data, bx, by, bz, rot);
  }
  return g;
}

FieldGrid read_srw(std::string const& file, Vec3 const& rot, char comment) {
  std::ifstream f(file);
  if (!f) throw std::invalid_argument("cannot open file for reading SRW format");
  next_line(f);
  FieldGrid g;
  for (int a = 0; a < 3; ++a) {
    g.start[a] = header_value_srw(next_line(f), comment);
    g.step[a] = header_value_srw(next_line(f), comment);
    g.n[a] = (int)header_value_srw(next_line(f), comment);
  }
  check_points(g.n[0], g.n[1], g.n[2]);
  finish_axes(g);
  size_t const np = (size_t)g.n[0] * g.n[1] * g.n[2];
  g.data.reserve(3 * np);
  std::string L;
  for (size_t k = 0; k < np; ++k) {
    std::getline(f, L);
    std::istringstream S(L);
    double fx, fy, fz;
    S >> fx >> fy >> fz;
    if (S.fail() || f.fail()) throw std::invalid_argument("input file stream failure");
    push_rotated(g.data, fx, fy, fz, rot);
  }
  return g;
}

// SPECTRA: "dx dy dz nx ny nz" with the steps in mm, the grid centred on the origin
// (TField3D_Grid.cc:1002-1146).
FieldGrid read_spectra(std::string const& file, Vec3 const& rot) {
  std::ifstream f(file);
  if (!f) throw std::invalid_argument("cannot open file");
  next_line(f);
  std::istringstream S(next_line(f));
  FieldGrid g;
  size_t n[3] = {0, 0, 0};
  S >> g.step[0] >> g.step[1] >> g.step[2] >> n[0] >> n[1] >> n[2];
  if (S.fail()) throw std::invalid_argument("file stream failure");
  for (int a = 0; a < 3; ++a) {
    g.step[a] /= 1000.;
    if (n[a] > (size_t)1 << 30) throw std::out_of_range("invalid number of points in at least one dimension");
    g.n[a] = (int)n[a];
  }
  check_points(g.n[0], g.n[1], g.n[2]);
  for (int a = 0; a < 3; ++a) g.start[a] = -(g.step[a] * (n[a] - 1)) / 2;
  for (int a = 0; a < 3; ++a) g.stop[a] = g.start[a] + (n[a] - 1) * g.step[a];
  size_t const np = n[0] * n[1] * n[2];
  g.data.reserve(3 * np);
  std::string L;
  for (size_t k = 0; k < np; ++k) {
    std::getline(f, L);
    if (f.eof()) throw std::invalid_argument("file stream failure");
    std::istringstream D(L);
    double fx, fy, fz;
    D >> fx >> fy >> fz;
    if (D.fail()) throw std::invalid_argument("file stream failure");
    push_rotated(g.data, fx, fy, fz, rot);
  }
  return g;
}

}  // namespace

FieldGrid read_grid_file(std::string const& file, std::string const& format, Vec3 const& rot,
                         std::vector<double> const& scale, char comment) {
  // OSCARSSR::AddMagneticField(file, ...), OSCARSSR.cc:68-90 and the TField3D_Grid
  // constructor (TField3D_Grid.cc:25-45): SRW and SPECTRA take no scaling
  std::string const F = upper(format);
  if (F == "OSCARS") return read_oscars(file, rot, scale);
  if (F == "SRW") return read_srw(file, rot, comment);
  if (F == "SPECTRA") return read_spectra(file, rot);
  if (F.size() > 8 && F.compare(0, 8, "OSCARS1D") == 0) return read_oscars1d(file, format, rot, scale, comment);
  throw std::invalid_argument("Incorrect format in format string");
}

FieldGrid read_grid_interpolated(std::vector<std::pair<double, std::string>> const& mapping,
                                 std::string const& format, double parameter, Vec3 const& rot,
                                 std::vector<double> const& scale) {
  // OSCARSSR::AddMagneticFieldInterpolated (OSCARSSR.cc:94-114) reads the OSCARS format
  // only (TField3D_Grid.cc:50-70: its other formats leave the grid unset); the maps are
  // sorted by parameter, must share one header, and each grid value is the linear
  // interpolation between the two maps that bracket the parameter
  // (InterpolateFromFiles / InterpolateFields, TField3D_Grid.cc:1150-1445).
  if (upper(format) != "OSCARS") throw std::invalid_argument("Incorrect format in format string");
  std::vector<std::pair<double, std::string>> m = mapping;
  std::sort(m.begin(), m.end(), [](std::pair<double, std::string> const& a,
                                          std::pair<double, std::string> const& b) { return a.first < b.first; });
  if (m.size() < 2) throw std::invalid_argument("need at least two maps to interpolate");
  int before = -1, after = -1;
  for (size_t i = 0; i != m.size(); ++i) {
    if (i == 0 && parameter < m[i].first) throw std::out_of_range("parameter below the mapping range");
    if (m[i].first < parameter) {
      before = (int)i;
    } else {
      after = (int)i;
      break;
    }
  }
  if (after < 0) throw std::out_of_range("parameter above the mapping range");
  if (before < 0) before = after;  // parameter equal to the first map's: that map (the reference leaves the index unset)
  std::vector<std::unique_ptr<std::ifstream>> files;
  for (auto const& e : m) {
    files.emplace_back(new std::ifstream(e.second));
    if (!*files.back()) throw std::invalid_argument("cannot open file for reading");
  }
  double hv[10] = {0};
  for (int k = 0; k < 10; ++k) {
    std::string const L0 = next_line(*files[0]);
    hv[k] = k == 0 ? 0.0 : header_value(L0);
    for (size_t i = 1; i < files.size(); ++i) {
      std::string const Li = next_line(*files[i]);
      if (k > 0 && header_value(Li) != hv[k]) throw std::invalid_argument("maps differ in their headers");
    }
  }
  FieldGrid g;
  oscars_axes(g, hv, scale);
  double const sx = scale.size() > 3 ? scale[3] : 1, sy = scale.size() > 4 ? scale[4] : 1,
               sz = scale.size() > 5 ? scale[5] : 1;
  size_t const np = (size_t)g.n[0] * g.n[1] * g.n[2];
  g.data.reserve(3 * np);
  std::vector<Vec3> F(files.size());
  std::string L;
  for (size_t k = 0; k < np; ++k) {
    for (size_t i = 0; i < files.size(); ++i) {
      std::getline(*files[i], L);
      if (files[i]->eof()) throw std::invalid_argument("error reading file.  Check format");
      std::istringstream S(L);
      double fx, fy, fz;
      S >> fx >> fy >> fz;
      if (sx != 1) fx *= sx;
      if (sy != 1) fy *= sy;
      if (sz != 1) fz *= sz;
      if (S.fail()) throw std::invalid_argument("error reading file.  Check format");
      F[i] = Vec3(fx, fy, fz);
    }
    double const diff = parameter - m[before].first;
    double const dp = m[after].first - m[before].first;
    Vec3 v = F[before];
    if (after != before) {
      Vec3 const slope((F[after].x - F[before].x) / dp, (F[after].y - F[before].y) / dp,
                       (F[after].z - F[before].z) / dp);
      v = Vec3(F[before].x + slope.x * diff, F[before].y + slope.y * diff, F[before].z + slope.z * diff);
    }
    push_rotated(g.data, v.x, v.y, v.z, rot);
  }
  return g;
}

Vec3 FieldContainer::eval(Vec3 const& X) const {
  Vec3 S(0, 0, 0);
  for (auto const& f : fields_) {
    Vec3 const b = f.eval(X);
    S.x += b.x;
    S.y += b.y;
    S.z += b.z;
  }
  return S;
}

// ---- particles -------------------------------------------------------------------

void Trajectory::clear() {
  for (auto* v : {&x, &y, &z, &bx, &by, &bz, &ax, &ay, &az}) v->clear();
  dt = 0;
}

void Particle::set_type(std::string const& t) {
  type = t;
  std::string const l = lower(t);
  if (l == "electron" || l == "anti-positron") {
    q = -kQe; m = kMe;
  } else if (l == "positron" || l == "anti-electron") {
    q = kQe; m = kMe;
  } else if (l == "muon") {
    q = -kQe; m = gev_to_kg(0.1056583715);
  } else if (l == "anti-muon") {
    q = kQe; m = gev_to_kg(0.1056583715);
  } else if (l == "proton") {
    q = kQe; m = gev_to_kg(0.938272046);
  } else if (l == "anti-proton") {
    q = -kQe; m = gev_to_kg(0.938272046);
  } else if (l == "pi+") {
    q = kQe; m = gev_to_kg(0.13957018);
  } else if (l == "pi-") {
    q = -kQe; m = gev_to_kg(0.13957018);
  } else if (l == "custom") {
  } else {
    throw std::invalid_argument("particle type not found: " + t);
  }
}

void Particle::set_current(double c) { current = c != 0 ? c : q; }

void Particle::set_initial(Vec3 const& X, Vec3 const& B, double T) {
  x0 = X;
  beta0 = B;
  t0 = T;
  double const b2 = mag2(beta0);
  if (b2 == 1) return;
  gamma = 1.0 / std::sqrt(1.0 - b2);
  if (m == 0 || gamma == 0) return;
  qmg = q / m / gamma;
}

double Particle::e0_gev() const { return kg_to_gev(m) * gamma; }

Particle ParticleBeam::ideal() const {
  double const g = e0 / kg_to_gev(proto.m);
  double const b = std::sqrt(1.0 - 1.0 / (g * g));
  Particle p = proto;
  p.set_initial(x0, scale(u0, b), t0);
  return p;
}

Particle ParticleBeam::random() const {
  Random& R = Random::global();
  double const enew = e0 + sigma_e * R.normal();
  double const g = enew / kg_to_gev(proto.m);
  double const b = std::sqrt(1.0 - 1.0 / (g * g));
  double const dmid = dot(sub(sigma_at, x0), u0);
  Vec3 X = x0;
  X = vadd(X, scale(scale(scale(hdir, sigma_u[0]), 1 + dmid), R.normal()));
  X = vadd(X, scale(scale(scale(vdir, sigma_u[1]), 1 + dmid), R.normal()));
  Vec3 B = scale(u0, b);
  rotate_y(B, sigma_up[0] * R.normal());
  rotate_x(B, -sigma_up[1] * R.normal());
  Particle p = proto;
  p.traj.clear();
  p.set_initial(X, B, t0);
  return p;
}

void ParticleBeam::set_sigma(Vec3 const& h, const double su[2], const double sup[2], Vec3 const& at, double se) {
  hdir = unit(h);
  sigma_u[0] = su[0];
  sigma_u[1] = su[1];
  sigma_up[0] = sup[0];
  sigma_up[1] = sup[1];
  sigma_at = at;
  sigma_e = se;
  vdir = unit(cross(u0, hdir));
  if (std::fabs(dot(hdir, u0)) > 0.00000001)
    throw std::invalid_argument("horizontal_direction must be orthogonal to the beam direction");
}

Random::Random() {
  std::random_device rd;
  mt_.reset(new std::mt19937(rd()));
}
Random& Random::global() {
  static Random r;
  return r;
}
void Random::seed(int s) { mt_.reset(new std::mt19937(s)); }
double Random::normal() { return normal_(*mt_); }
double Random::uniform() { return uniform_(*mt_); }

// ---- surfaces and containers ---------------------------------------------------

SurfacePoints SurfacePoints::rectangle(std::string const& plane, int nx1, int nx2, double w1, double w2, Vec3 const& rot,
                                       Vec3 const& trans, int normal) {
  // TSurfacePoints_Rectangle::Init(plane, ...), .cc:130-210; GetXYZ :222-230; GetX1/X2 :244-259
  std::string const P = upper(plane);
  double const s1 = w1 / (nx1 - 1), s2 = w2 / (nx2 - 1);
  Vec3 st, a1, a2;
  if (P == "XY") {
    st = Vec3(-w1 / 2., -w2 / 2., 0); a1 = Vec3(s1, 0, 0); a2 = Vec3(0, s2, 0);
  } else if (P == "YX") {
    st = Vec3(-w2 / 2., -w1 / 2., 0); a2 = Vec3(s2, 0, 0); a1 = Vec3(0, s1, 0);
  } else if (P == "XZ") {
    st = Vec3(-w1 / 2., 0, -w2 / 2.); a1 = Vec3(s1, 0, 0); a2 = Vec3(0, 0, s2);
  } else if (P == "ZX") {
    st = Vec3(-w2 / 2., 0, -w1 / 2.); a2 = Vec3(s2, 0, 0); a1 = Vec3(0, 0, s1);
  } else if (P == "YZ") {
    st = Vec3(0, -w1 / 2., -w2 / 2.); a1 = Vec3(0, s1, 0); a2 = Vec3(0, 0, s2);
  } else if (P == "ZY") {
    st = Vec3(0, -w2 / 2., -w1 / 2.); a2 = Vec3(0, s2, 0); a1 = Vec3(0, 0, s1);
  } else {
    throw std::invalid_argument("not a valid surface string: XY YX XZ ZX YZ ZY");
  }
  rotate_xyz(st, rot);
  rotate_xyz(a1, rot);
  rotate_xyz(a2, rot);
  st = vadd(st, trans);
  Vec3 nv = unit(cross(a1, a2));
  if (normal == -1) nv = scale(nv, -1);
  else if (normal != 0 && normal != 1) throw std::invalid_argument("normal must be -1, 0, or 1");
  SurfacePoints S;
  S.has_x12 = true;
  size_t const n = (size_t)nx1 * (size_t)nx2;
  S.xyz.resize(3 * n);
  S.normals.resize(3 * n);
  S.x1.resize(n);
  S.x2.resize(n);
  double const m1 = mag(a1), m2 = mag(a2);
  for (size_t i = 0; i < n; ++i) {
    int const i1 = (int)(i / nx2), i2 = (int)(i % nx2);
    Vec3 const p = vadd(vadd(st, scale(a1, (double)i1)), scale(a2, (double)i2));
    S.xyz[3 * i] = p.x;
    S.xyz[3 * i + 1] = p.y;
    S.xyz[3 * i + 2] = p.z;
    S.normals[3 * i] = nv.x;
    S.normals[3 * i + 1] = nv.y;
    S.normals[3 * i + 2] = nv.z;
    S.x1[i] = s1 * i1 - m1 * (nx1 - 1) / 2.;
    S.x2[i] = s2 * i2 - m2 * (nx2 - 1) / 2.;
  }
  return S;
}

SurfacePoints SurfacePoints::rectangle3(int nx1, int nx2, Vec3 const& X0, Vec3 const& X1, Vec3 const& X2, int normal) {
  // TSurfacePoints_Rectangle::Init(NX1, NX2, X0, X1, X2, Normal), .cc:76-123
  double const w1 = mag(sub(X1, X0)), w2 = mag(sub(X2, X0));
  double const s1 = w1 / (nx1 - 1), s2 = w2 / (nx2 - 1);
  Vec3 const a1 = divide(sub(X1, X0), (double)(nx1 - 1));
  Vec3 const a2 = divide(sub(X2, X0), (double)(nx2 - 1));
  Vec3 nv = unit(cross(a1, a2));
  if (normal == -1) nv = scale(nv, -1);
  else if (normal != 0 && normal != 1) throw std::invalid_argument("normal must be -1, 0, or 1");
  SurfacePoints S;
  S.has_x12 = true;
  size_t const n = (size_t)nx1 * (size_t)nx2;
  double const m1 = mag(a1), m2 = mag(a2);
  for (size_t i = 0; i < n; ++i) {
    int const i1 = (int)(i / nx2), i2 = (int)(i % nx2);
    Vec3 const p = vadd(vadd(X0, scale(a1, (double)i1)), scale(a2, (double)i2));
    S.add(p, nv);
    S.x1.push_back(s1 * i1 - m1 * (nx1 - 1) / 2.);
    S.x2.push_back(s2 * i2 - m2 * (nx2 - 1) / 2.);
  }
  return S;
}

void SurfacePoints::add(Vec3 const& X, Vec3 const& N) {
  xyz.insert(xyz.end(), {X.x, X.y, X.z});
  normals.insert(normals.end(), {N.x, N.y, N.z});
}

void SpectrumContainer::init(size_t n, double e0, double e1) {
  if (n < 1) throw std::length_error("'npoints' must be > 0");
  e.assign(n, 0.0);
  f.assign(n, 0.0);
  comp.assign(n, 0.0);
  if (n == 1) {
    e[0] = e0;
    return;
  }
  for (size_t i = 0; i != n; ++i) e[i] = e0 + (e1 - e0) / (n - 1) * (double)(i + 1);
}

void SpectrumContainer::init(std::vector<double> const& pts) {
  e = pts;
  f.assign(pts.size(), 0.0);
  comp.assign(pts.size(), 0.0);
}

void SpectrumContainer::add_to(size_t i, double v) {
  if (i >= f.size()) throw std::out_of_range("spectrum index out of range");
  double const s = f[i];
  double const y = v - comp[i];
  double const t = s + y;
  comp[i] = (t - s) - y;
  f[i] = t;
}

void ScalarContainer::add_point(Vec3 const& X, double val) {
  xyz.insert(xyz.end(), {X.x, X.y, X.z});
  v.push_back(val);
  comp.push_back(0);
}

void ScalarContainer::add_to(size_t i, double val) {
  if (i >= v.size()) throw std::length_error("T3DScalarContainer::AddtoPoint index out of range");
  double const s = v[i];
  double const y = val - comp[i];
  double const t = s + y;
  comp[i] = (t - s) - y;
  v[i] = t;
}

// ---- engine --------------------------------------------------------------------------

OSCARSSR::OSCARSSR() {}

void OSCARSSR::AddParticleBeam(std::string const& type, std::string const& name, Vec3 const& x0, Vec3 const& d0,
                               double e_gev, double t0, double current, double weight) {
  // TParticleBeamContainer::AddNewParticleBeam, .cc:35-58
  if (beam_index_.count(name)) throw std::invalid_argument("beam with this name already exists");
  ParticleBeam b;
  b.proto.set_type(type);
  b.proto.set_current(current);
  b.x0 = x0;
  b.u0 = unit(d0);
  b.e0 = e_gev;
  b.t0 = t0;
  b.weight = weight;
  beam_weight_sums_.push_back(beam_weight_sums_.empty() ? weight : beam_weight_sums_.back() + weight);
  beams_.push_back(b);
  beam_index_[name] = beams_.size() - 1;
}

ParticleBeam& OSCARSSR::GetParticleBeam(std::string const& name) {
  auto it = beam_index_.find(name);
  if (it == beam_index_.end()) throw std::out_of_range("beam name not in map");
  return beams_[it->second];
}

void OSCARSSR::ClearParticleBeams() {
  beams_.clear();
  beam_weight_sums_.clear();
  beam_index_.clear();
}

namespace {
size_t beam_by_weight(std::vector<double> const& sums) {
  // TParticleBeamContainer::GetRandomBeamIndexByWeight, .cc:109-142
  size_t const n = sums.size();
  if (n == 0) throw std::length_error("no beam defined");
  if (n == 1) return 0;
  double const r = Random::global().uniform() * sums[n - 1];
  for (size_t i = 0; i != n; ++i)
    if (r < sums[i]) return i;
  throw std::out_of_range("random weight out of range");
}
}  // namespace

void OSCARSSR::SetNewParticle() { particle_ = beams_[beam_by_weight(beam_weight_sums_)].random(); }

void OSCARSSR::SetNewParticle(std::string const& beam, std::string const& ideal_or_random) {
  ParticleBeam& b = beam.empty() ? beams_[beam_by_weight(beam_weight_sums_)] : GetParticleBeam(beam);
  particle_ = lower(ideal_or_random) == "ideal" ? b.ideal() : b.random();
}

void OSCARSSR::SetNPointsPerMeter(size_t n) {
  npoints_per_meter_ = n;
  if (ct_start_ != ct_stop_) SetCTStartStop(ct_start_, ct_stop_);
}

void OSCARSSR::SetCTStartStop(double start, double stop) {
  // OSCARSSR.cc:448-458: also resets the number of points
  npoints_ = npoints_per_meter_ * (stop - start);
  ct_start_ = start;
  ct_stop_ = stop;
}

int OSCARSSR::CheckGPU() const { return srs_hip_device_count(); }

void OSCARSSR::ensure_particle() {
  if (particle_.type.empty()) {
    if (beams_.empty()) throw std::out_of_range("no beam defined");
    SetNewParticle();
  }
}

void OSCARSSR::CalculateTrajectory() {
  if (particle_.type.empty()) SetNewParticle();
  CalculateTrajectory(particle_);
}

void OSCARSSR::CalculateTrajectory(Particle& P) const {
  // OSCARSSR::CalculateTrajectory(TParticleA&), OSCARSSR.cc:562-673.  Forward pass
  // stores the state before each step with the derivative of the previous step
  // (A/c lags one point; the first point's A/c, uninitialised in the reference, is 0);
  // the backward pass fills points before T0.
  if (ct_start_ > P.t0) throw std::out_of_range("start time is greater than T0");
  if (ct_start_ >= ct_stop_) throw std::out_of_range("start time is greater than stop time.  check that set_ctstartstop is set");
  if (P.type.empty()) throw std::out_of_range("particle not initialized.  make sure you have a particle or beam defined");
  if (npoints_ < 2) throw std::out_of_range("number of trajectory points must be >= 2");
  Trajectory& T = P.traj;
  T.clear();
  double const dt = ((ct_stop_ - ct_start_) / kC / (npoints_ - 1));
  size_t const nf = 1 + (ct_stop_ - P.t0) / kC / dt;
  size_t const nb = (P.t0 - ct_start_) / kC / dt;
  size_t const n = nf + nb;
  T.dt = dt;
  for (auto* v : {&T.x, &T.y, &T.z, &T.bx, &T.by, &T.bz, &T.ax, &T.ay, &T.az}) v->assign(n, 0.0);
  double const qmg = P.qmg;
  FieldContainer const& F = bfield_;
  FieldContainer const& G = efield_;
  // OSCARSSR::SetDerivativesFunction (OSCARSSR.cc:690-703): DerivativesE (:713-737) when
  // only electric fields exist, DerivativesB (:741-765) when only magnetic ones, else
  // DerivativesEB (:769-794).  The three agree in value where a field is absent but not
  // in the sign of zero results, which the stored A/c keeps.
  bool const has_b = !F.fields().empty(), has_e = !G.fields().empty();
  auto derivs = [&](const double* x, double* d) {
    Vec3 const X(x[0], x[2], x[4]);
    d[0] = x[1];
    d[2] = x[3];
    d[4] = x[5];
    if (has_e && !has_b) {
      Vec3 const E = G.eval(X);
      d[1] = qmg * E.x;
      d[3] = qmg * E.y;
      d[5] = qmg * E.z;
    } else if (has_b && !has_e) {
      Vec3 const B = F.eval(X);
      d[1] = qmg * (-x[5] * B.y + x[3] * B.z);
      d[3] = qmg * (x[5] * B.x - x[1] * B.z);
      d[5] = qmg * (x[1] * B.y - x[3] * B.x);
    } else {
      Vec3 const B = F.eval(X);
      Vec3 const E = G.eval(X);
      d[1] = qmg * (E.x - x[5] * B.y + x[3] * B.z);
      d[3] = qmg * (E.y + x[5] * B.x - x[1] * B.z);
      d[5] = qmg * (E.z + x[1] * B.y - x[3] * B.x);
    }
  };
  auto rk4 = [&](double* y, const double* dydx, double h) {
    double dym[6], dyt[6], yt[6];
    double const hh = h * 0.5, h6 = h / 6.0;
    for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dydx[i];
    derivs(yt, dyt);
    for (int i = 0; i < 6; ++i) yt[i] = y[i] + hh * dyt[i];
    derivs(yt, dym);
    for (int i = 0; i < 6; ++i) {
      yt[i] = y[i] + h * dym[i];
      dym[i] += dyt[i];
    }
    derivs(yt, dyt);
    for (int i = 0; i < 6; ++i) y[i] = y[i] + h6 * (dydx[i] + dyt[i] + 2.0 * dym[i]);
  };
  double x[6], d[6] = {0, 0, 0, 0, 0, 0};
  auto init = [&]() {
    x[0] = P.x0.x;
    x[1] = P.beta0.x * kC;
    x[2] = P.x0.y;
    x[3] = P.beta0.y * kC;
    x[4] = P.x0.z;
    x[5] = P.beta0.z * kC;
  };
  auto store = [&](size_t k) {
    T.x[k] = x[0];
    T.y[k] = x[2];
    T.z[k] = x[4];
    T.bx[k] = x[1] / kC;
    T.by[k] = x[3] / kC;
    T.bz[k] = x[5] / kC;
    T.ax[k] = d[1] / kC;
    T.ay[k] = d[3] / kC;
    T.az[k] = d[5] / kC;
  };
  init();
  for (size_t i = 0; i != nf; ++i) {
    store(nb + i);
    derivs(x, d);
    rk4(x, d, dt);
  }
  init();
  double const mdt = -dt;
  for (size_t i = 0; i != nb; ++i) {
    derivs(x, d);
    rk4(x, d, mdt);
    store(nb - 1 - i);
  }
}

// ---- radiation ---------------------------------------------------------------------------

void OSCARSSR::spectrum_one(Particle& p, Vec3 const& obs, SpectrumContainer& S, double weight, Precision prec) {
  CalculateTrajectory(p);
  srs_traj const t = view(p.traj);
  double const o[3] = {obs.x, obs.y, obs.z};
  std::vector<double> out(S.e.size());
  check_status(srs_hip_spectrum(&t, o, S.e.data(), (int64_t)S.e.size(), p.q, p.current, (int)prec, out.data()));
  for (size_t i = 0; i < out.size(); ++i) S.add_to(i, out[i] * weight);
}

void OSCARSSR::CalculateSpectrum(Vec3 const& obs, SpectrumContainer& S, int nparticles, int nthreads, int gpu,
                                 Precision prec, bool batched) {
  // OSCARSSR::CalculateSpectrum(Obs, S, NParticles, NThreads, GPU), OSCARSSR.cc:933-984.
  // Every route accumulates value * weight; the integrals run on the GPU.
  (void)nthreads;
  ensure_particle();
  if (nparticles == 0) {
    spectrum_one(particle_, obs, S, 1, prec);
    return;
  }
  double const w = 1.0 / (double)nparticles;
  if (batched && gpu == 1 && nparticles > 1) {
    std::vector<srs_particle> parts;
    parts.reserve(nparticles);
    double q = 0, cur = 0;
    for (int i = 0; i != nparticles; ++i) {
      SetNewParticle();
      parts.push_back(device_particle(particle_));
      if (i == 0) {
        q = particle_.q;  // the call's defaults; every particle carries its own
        cur = particle_.current;
      }
    }
    if (no_fields()) throw std::invalid_argument("no magnetic or electric field given");
    std::vector<srs_field> und;
    device_fields(und);
    double const o[3] = {obs.x, obs.y, obs.z};
    std::vector<double> out(S.e.size());
    check_status(srs_hip_batch_spectrum(und.data(), (int)und.size(), parts.data(), nparticles, q, cur, o, S.e.data(),
                                        (int64_t)S.e.size(), w, (int)prec, out.data()));
    for (size_t i = 0; i < out.size(); ++i) S.add_to(i, out[i]);
    return;
  }
  for (int i = 0; i != nparticles; ++i) {
    SetNewParticle();
    spectrum_one(particle_, obs, S, w, prec);
  }
}

namespace {
void add_points(SurfacePoints const& surf, ScalarContainer& C, int dim) {
  size_t const n = surf.size();
  for (size_t i = 0; i < n; ++i) {
    if (dim == 3) {
      C.add_point(Vec3(surf.xyz[3 * i], surf.xyz[3 * i + 1], surf.xyz[3 * i + 2]), 0);
    } else if (dim == 2) {
      C.add_point(surf.has_x12 ? Vec3(surf.x1[i], surf.x2[i], 0) : Vec3(0, 0, 0), 0);
    } else {
      throw std::out_of_range("'dim' must be 2 or 3");
    }
  }
}
}  // namespace

void OSCARSSR::device_fields(std::vector<srs_field>& out) const {
  out.clear();
  for (int target = 0; target < 2; ++target)
  for (auto const& u : (target == 0 ? bfield_ : efield_).fields()) {
    srs_field d{};
    d.target = target == 0 ? SRS_FIELD_MAGNETIC : SRS_FIELD_ELECTRIC;
    d.field[0] = u.field.x;
    d.field[1] = u.field.y;
    d.field[2] = u.field.z;
    d.period[0] = u.period.x;
    d.period[1] = u.period.y;
    d.period[2] = u.period.z;
    d.nperiods = u.nperiods;
    d.phase = u.phase;
    d.taper = u.taper;
    d.center[0] = u.center.x;
    d.center[1] = u.center.y;
    d.center[2] = u.center.z;
    d.pole_delta = u.pole_delta.empty() ? nullptr : u.pole_delta.data();
    d.kind = (int32_t)u.kind;
    d.width[0] = u.width.x;
    d.width[1] = u.width.y;
    d.width[2] = u.width.z;
    d.rotations[0] = u.rot.x;
    d.rotations[1] = u.rot.y;
    d.rotations[2] = u.rot.z;
    if (u.grid) {
      d.grid = u.grid->data.data();
      for (int k = 0; k < 3; ++k) {
        d.grid_n[k] = u.grid->n[k];
        d.grid_start[k] = u.grid->start[k];
        d.grid_step[k] = u.grid->step[k];
        d.grid_stop[k] = u.grid->stop[k];
      }
    }
    out.push_back(d);
  }
}

srs_particle OSCARSSR::device_particle(Particle const& p) const {
  srs_particle d{};
  d.x0[0] = p.x0.x;
  d.x0[1] = p.x0.y;
  d.x0[2] = p.x0.z;
  d.beta0[0] = p.beta0.x;
  d.beta0[1] = p.beta0.y;
  d.beta0[2] = p.beta0.z;
  d.qmg = p.qmg;
  d.t0 = p.t0;
  d.ct_start = ct_start_;
  d.ct_stop = ct_stop_;
  d.charge = p.q;  // each particle's own prefactors, as the reference's per-particle calls have them
  d.current = p.current;
  d.npoints = (int64_t)npoints_;
  return d;
}

void OSCARSSR::CalculateTrajectoryGPU() {
  if (particle_.type.empty()) SetNewParticle();
  CalculateTrajectoryGPU(particle_);
}

void OSCARSSR::CalculateTrajectoryGPU(Particle& P) const {
  // Same checks and point layout as CalculateTrajectory (OSCARSSR.cc:562-673).
  if (ct_start_ > P.t0) throw std::out_of_range("start time is greater than T0");
  if (ct_start_ >= ct_stop_) throw std::out_of_range("start time is greater than stop time.  check that set_ctstartstop is set");
  if (P.type.empty()) throw std::out_of_range("particle not initialized.  make sure you have a particle or beam defined");
  if (npoints_ < 2) throw std::out_of_range("number of trajectory points must be >= 2");
  if (no_fields()) throw std::invalid_argument("no magnetic or electric field given");
  std::vector<srs_field> und;
  device_fields(und);
  srs_particle const dp = device_particle(P);
  int64_t total = 0;
  check_status(srs_hip_trajectory_points(&dp, 1, &total));
  std::vector<double> soa((size_t)9 * total);
  int64_t offs[2];
  double dt = 0;
  check_status(srs_hip_trajectories(und.data(), (int)und.size(), &dp, 1, soa.data(), total, offs, &dt));
  Trajectory& T = P.traj;
  T.clear();
  T.dt = dt;
  std::vector<double>* dst[9] = {&T.x, &T.y, &T.z, &T.bx, &T.by, &T.bz, &T.ax, &T.ay, &T.az};
  for (int k = 0; k < 9; ++k) dst[k]->assign(soa.begin() + (size_t)k * total, soa.begin() + (size_t)(k + 1) * total);
}

void OSCARSSR::flux_one(Particle& p, SurfacePoints const& surf, double energy_ev, ScalarContainer& C, double weight,
                        Precision prec) {
  CalculateTrajectory(p);
  srs_traj const t = view(p.traj);
  std::vector<double> out(surf.size());
  check_status(srs_hip_flux(&t, surf.xyz.data(), (int64_t)surf.size(), &energy_ev, 1, p.q, p.current, (int)prec,
                            out.data()));
  for (size_t i = 0; i < out.size(); ++i) C.add_to(i, out[i] * weight);
}

void OSCARSSR::CalculateFlux(SurfacePoints const& surf, double energy_ev, ScalarContainer& C, int nparticles,
                             int nthreads, int gpu, int dim, Precision prec, bool batched) {
  // OSCARSSR::CalculateFlux(Surface, E, C, NP, NT, GPU, Dim), OSCARSSR.cc:2418-2489.
  // Container layout follows the reference route (OSCARSSR.cc:2352-2358, :2512, :2570-2572):
  //   gpu=1: N extra 3-D points per particle; gpu=0 with 1 thread: N extra points per
  //   particle; gpu=0 with threads: none.  Values land in the first N entries.
  int const nt = nthreads < 1 ? nthreads_global_ : nthreads;
  if (nt <= 0) throw std::out_of_range("NThreads or NThreadsGlobal must be >= 1");
  ensure_particle();
  add_points(surf, C, dim);
  int const np = nparticles == 0 ? 1 : nparticles;
  double const w = nparticles == 0 ? 1.0 : 1.0 / (double)nparticles;
  if (batched && gpu == 1 && nparticles > 1) {
    // draw all particles first (same draws, same order), then one device pipeline
    std::vector<srs_particle> parts;
    parts.reserve(np);
    double q = 0, cur = 0;
    for (int i = 0; i != np; ++i) {
      SetNewParticle();
      add_points(surf, C, 3);
      parts.push_back(device_particle(particle_));
      if (i == 0) {
        q = particle_.q;  // the call's defaults; every particle carries its own
        cur = particle_.current;
      }
    }
    if (no_fields()) throw std::invalid_argument("no magnetic or electric field given");
    std::vector<srs_field> und;
    device_fields(und);
    std::vector<double> out(surf.size());
    check_status(srs_hip_batch_flux(und.data(), (int)und.size(), parts.data(), np, q, cur, surf.xyz.data(),
                                    (int64_t)surf.size(), &energy_ev, 1, w, (int)prec, SRS_BATCH_GRAPH, out.data()));
    for (size_t i = 0; i < out.size(); ++i) C.add_to(i, out[i]);
    return;
  }
  for (int i = 0; i != np; ++i) {
    if (nparticles != 0) SetNewParticle();
    if (gpu == 1) {
      add_points(surf, C, 3);
    } else if (nt == 1) {
      add_points(surf, C, dim);
    }
    flux_one(particle_, surf, energy_ev, C, w, prec);
  }
}

void OSCARSSR::pd_one(Particle& p, SurfacePoints const& surf, ScalarContainer& C, bool directional, double weight,
                      Precision prec) {
  CalculateTrajectory(p);
  srs_traj const t = view(p.traj);
  std::vector<double> out(surf.size());
  check_status(srs_hip_power_density(&t, surf.xyz.data(), surf.normals.data(), (int64_t)surf.size(),
                                     directional ? 1 : 0, p.q, p.current, (int)prec, out.data()));
  for (size_t i = 0; i < out.size(); ++i) C.add_to(i, out[i] * weight);
}

void OSCARSSR::CalculatePowerDensity(SurfacePoints const& surf, ScalarContainer& C, int dim, bool directional,
                                     int nparticles, int nthreads, int gpu, Precision prec, bool batched) {
  // OSCARSSR::CalculatePowerDensity(Surface, C, Dim, Dir, NP, ofile, NT, GPU), OSCARSSR.cc:1603-1677.
  // Reference route semantics kept: gpu=0 ignores the particle weight (sum over
  // particles, OSCARSSR.cc:1579-1585) and its threads route adds N extra points
  // (:1804-1811); gpu=1 applies the weight and adds N extra 3-D points (:1869-1871).
  int const nt = nthreads < 1 ? nthreads_global_ : nthreads;
  if (nt <= 0) throw std::out_of_range("NThreads or NThreadsGlobal must be >= 1");
  ensure_particle();
  add_points(surf, C, dim);
  int const np = nparticles == 0 ? 1 : nparticles;
  double const w = nparticles == 0 ? 1.0 : 1.0 / (double)nparticles;
  if (batched && gpu == 1 && nparticles > 1) {
    // the gpu=1 route's draws and container points, then one device pipeline
    std::vector<srs_particle> parts;
    parts.reserve(np);
    double q = 0, cur = 0;
    for (int i = 0; i != np; ++i) {
      SetNewParticle();
      add_points(surf, C, 3);
      parts.push_back(device_particle(particle_));
      if (i == 0) {
        q = particle_.q;  // the call's defaults; every particle carries its own
        cur = particle_.current;
      }
    }
    if (no_fields()) throw std::invalid_argument("no magnetic or electric field given");
    std::vector<srs_field> und;
    device_fields(und);
    std::vector<double> out(surf.size());
    check_status(srs_hip_batch_power_density(und.data(), (int)und.size(), parts.data(), np, q, cur, surf.xyz.data(),
                                             surf.normals.data(), (int64_t)surf.size(), directional ? 1 : 0, w,
                                             (int)prec, out.data()));
    for (size_t i = 0; i < out.size(); ++i) C.add_to(i, out[i]);
    return;
  }
  for (int i = 0; i != np; ++i) {
    if (nparticles != 0) SetNewParticle();
    if (gpu == 1) {
      add_points(surf, C, 3);
      pd_one(particle_, surf, C, directional, w, prec);
    } else {
      if (nt != 1) add_points(surf, C, dim);
      pd_one(particle_, surf, C, directional, 1.0, prec);
    }
  }
}

void OSCARSSR::CalculateElectricFieldTimeDomain(Vec3 const& obs, std::vector<double>& e, std::vector<double>& t) {
  ensure_particle();
  CalculateTrajectory(particle_);
  srs_traj const tr = view(particle_.traj);
  double const o[3] = {obs.x, obs.y, obs.z};
  e.assign(3 * particle_.traj.size(), 0.0);
  t.assign(particle_.traj.size(), 0.0);
  check_status(srs_hip_efield_time(&tr, o, particle_.q, e.data(), t.data()));
}

void OSCARSSR::CalculateFlux1(Particle& P, SurfacePoints const& surf, double energy_ev, ScalarContainer& C,
                              std::string const& ofile) {
  // OSCARSSR.cc:2209-2346: particle check, trajectory, the sum per surface point (on the
  // device), then AddToPoint(ip, flux) or one "X1 X2 flux" line per point (Dimension is
  // hard-coded to 2 there; point lists have X1 = X2 = 0).
  if (P.type.empty()) throw std::out_of_range("no particle defined");
  std::ofstream of;
  if (!ofile.empty()) {
    of.open(ofile.c_str());
    if (!of.is_open()) throw std::ofstream::failure("cannot open output file");
    of << std::scientific;
  }
  CalculateTrajectory(P);
  srs_traj const t = view(P.traj);
  std::vector<double> out(surf.size());
  if (!out.empty())
    check_status(srs_hip_flux1(&t, surf.xyz.data(), (int64_t)surf.size(), energy_ev, P.q, P.current, out.data()));
  for (size_t i = 0; i < out.size(); ++i) {
    if (of.is_open()) {
      of << (surf.has_x12 ? surf.x1[i] : 0.0) << " " << (surf.has_x12 ? surf.x2[i] : 0.0) << " " << out[i] << "\n";
    } else {
      C.add_to(i, out[i]);
    }
  }
}

void OSCARSSR::CalculateFlux1(SurfacePoints const& surf, double energy_ev, ScalarContainer& C,
                              std::string const& ofile) {
  ensure_particle();
  CalculateFlux1(particle_, surf, energy_ev, C, ofile);
}

double OSCARSSR::CalculateTotalPower() {
  // OSCARSSR::CalculateTotalPower(TParticleA&), OSCARSSR.cc:1939-1978 (host, O(NT))
  ensure_particle();
  CalculateTrajectory(particle_);
  Trajectory const& T = particle_.traj;
  double tp = 0;
  for (size_t i = 0; i != T.size(); ++i) {
    Vec3 const B(T.bx[i], T.by[i], T.bz[i]);
    Vec3 const A(T.ax[i], T.ay[i], T.az[i]);
    tp += (mag2(A) - mag2(cross(B, A))) * T.dt;
  }
  const double kEps0 = 8.854187817E-12;
  tp *= std::fabs(particle_.q * particle_.current) * std::pow(particle_.gamma, 6) / (6 * kPi * kEps0 * kC);
  return tp;
}

double OSCARSSR::CalculateTotalPowerDevice() {
  // the same quantity, the trajectory from the host RK4 and the sum on the device
  ensure_particle();
  CalculateTrajectory(particle_);
  srs_traj const v = view(particle_.traj);
  double p = 0;
  check_status(srs_hip_total_power(&v, particle_.q, particle_.current, particle_.gamma, &p));
  return p;
}

}  // namespace srs
