// zpulse.hip -- laser injection from an x wall (zpulse_wall), 2D and 3D: host-side set-up
// and the FP64 evaluation of the column (plane) a pulse adds to B2, B3 at one time.
//
// Reference: zpulse/os-zpulse-wall.f03 (read_input_zpulse_wall :90-628, launch_zpulse_wall
// :633-700, t_env_zpulse_wall :705-746, t_duration :752-771, lon_center :781-805,
// launch_em_wall_2d :865-960, launch_em_wall_3d :965-1082), per_envelope_2d_zpulse /
// per_envelope_3d_zpulse zpulse/os-zpulse-std.f03:2599-2753, 2760-3049, fenv_poly /
// fenv_sin2 os-math.f90:83-165, hermite_double os-math.f90:478-509.
//
// The evaluation keeps the reference's operation order in double precision; the values are
// rounded to float (real(., p_k_fld)) and the device adds them to B in FP32.  Included
// by capi.hip.
namespace osr_zp {

constexpr double kPi2 = 1.57079632679489661923132;     // os-math.f90:14
constexpr double kPi180 = 0.0174532925199432957692369;  // os-math.f90:15
constexpr int kMaxPulses = 8;

struct Wall {
  double a0, omega0, phase0, pol;  // phase0 and pol in radians
  int pol_type, backward, tenv_type, per_type, tem_mode[2];
  double rise, flat, fall, duration, range;
  double center[2], w0, focus;
  double launch_time;
  double xmin[3], xmax[3];
};

// os-math.f90:83-128
inline double fenv_poly(double x, double rise, double flat, double fall) {
  const double length = rise + flat + fall;
  auto f = [](double tt) { return 10 * (tt * tt * tt) - 15 * (tt * tt * tt * tt) + 6 * (tt * tt * tt * tt * tt); };
  if (x < 0.0) return 0.0;
  if (x < rise) return f(x / rise);
  if (x < rise + flat) return 1.0;
  if (x < length) return f((length - x) / fall);
  return 0.0;
}

// os-math.f90:132-165
inline double fenv_sin2(double x, double rise, double flat, double fall) {
  const double length = rise + flat + fall;
  if (x < 0.0) return 0.0;
  if (x < rise) {
    const double s = std::sin(kPi2 * x / rise);
    return s * s;
  }
  if (x < rise + flat) return 1.0;
  if (x < length) {
    const double s = std::sin(kPi2 * (length - x) / fall);
    return s * s;
  }
  return 0.0;
}

// os-math.f90:478-509
inline double hermite(int mode, double u) {
  switch (mode) {
    case 0: return 1.0;
    case 1: return 2.0 * u;
    case 2: return 4.0 * (u * u) - 2.0;
    case 3: return 8.0 * (u * u * u) - 12.0 * u;
    case 4: return 16.0 * (u * u * u * u) - 48.0 * (u * u) + 12.0;
    case 5: return 32.0 * (u * u * u * u * u) - 160.0 * (u * u * u) + 120.0 * u;
    case 6: return 64.0 * (u * u * u * u * u * u) - 480.0 * (u * u * u * u) + 720.0 * (u * u) - 120.0;
    case 7:
      return 128.0 * (u * u * u * u * u * u * u) - 1344.0 * (u * u * u * u * u) +
             3360.0 * (u * u * u) - 1680.0 * u;
    default: return 0.0;
  }
}

inline double t_duration(const Wall& w) {
  return w.tenv_type == OSR_TENV_GAUSSIAN ? w.range : w.rise + w.flat + w.fall;
}

inline double t_envelope(const Wall& w, double t) {
  switch (w.tenv_type) {
    case OSR_TENV_POLYNOMIAL: return fenv_poly(t, w.rise, w.flat, w.fall);
    case OSR_TENV_SIN2: return fenv_sin2(t, w.rise, w.flat, w.fall);
    default: {
      const double a = (t - 0.5 * w.range) / w.duration;
      return std::exp(-2 * (a * a));
    }
  }
}

// per_envelope_2d_zpulse at simulation time 0 (the wall calls it without a time); the
// pulse centre moves with the injection time through lon_center
inline double per_envelope_2d(const Wall& w, double x1, double x2, double inj_time) {
  const double lon = w.tenv_type == OSR_TENV_GAUSSIAN ? 0.5 * w.range : w.rise + w.flat / 2;
  const double z_center = lon + inj_time;
  const double k = w.omega0;
  const double prop_sign = w.backward ? -1.0 : 1.0;
  const double t = 0.0;
  if (w.per_type == OSR_PER_PLANE) return std::cos(k * (x1 - z_center - prop_sign * t) + w.phase0);
  const double z0 = k * (w.w0 * w.w0) * 0.5;
  const double rho = x2 - w.center[0];
  const double rho2 = rho * rho;
  const double z = x1 - w.focus;
  double rWl2 = 1, curv = 0, gouy = 0;
  if (z != 0) {
    rWl2 = (z0 * z0) / ((z0 * z0) + (z * z));
    curv = 0.5 * rho2 * z / ((z * z) + (z0 * z0));
    gouy = (w.tem_mode[0] + 1) * std::atan2(z, z0);
  }
  const double rWu = std::sqrt(2 * rWl2) / w.w0;
  return std::sqrt(std::sqrt(rWl2)) * hermite(w.tem_mode[0], rho * rWu) *
         std::exp(-(rho2 * rWl2 / (w.w0 * w.w0))) *
         std::cos(k * (x1 - z_center - prop_sign * t) + k * curv - gouy + w.phase0);
}

// per_envelope_3d_zpulse (plane, hermite_gaussian) at simulation time 0
inline double per_envelope_3d(const Wall& w, double x1, double x2, double x3, double inj_time) {
  const double lon = w.tenv_type == OSR_TENV_GAUSSIAN ? 0.5 * w.range : w.rise + w.flat / 2;
  const double z_center = lon + inj_time;
  const double ox = x2 - w.center[0], oy = x3 - w.center[1];
  const double k = w.omega0;
  const double prop_sign = w.backward ? -1.0 : 1.0;
  const double t = 0.0;
  if (w.per_type == OSR_PER_PLANE) return std::cos(k * (x1 - z_center - prop_sign * t) + w.phase0);
  const double rho2 = (ox * ox) + (oy * oy);
  const double z = x1 - w.focus;
  const double z0 = k * (w.w0 * w.w0) / 2;
  double rWl2 = 1, curv = 0, gouy = 0;
  if (z != 0) {
    rWl2 = (z0 * z0) / ((z0 * z0) + (z * z));
    curv = 0.5 * rho2 * z / ((z * z) + (z0 * z0));
    gouy = (w.tem_mode[0] + w.tem_mode[1] + 1) * std::atan2(z, z0);
  }
  const double rWu = std::sqrt(2 * rWl2) / w.w0;
  return std::sqrt(rWl2) * hermite(w.tem_mode[0], ox * rWu) * hermite(w.tem_mode[1], oy * rWu) *
         std::exp(-(rho2 * rWl2 / (w.w0 * w.w0))) *
         std::cos(k * (x1 - z_center - prop_sign * t) + k * curv - gouy + w.phase0);
}

// Whether the pulse adds anything at time t (launch_zpulse_wall: active from launch_time
// through launch_time + duration).
inline bool active(const Wall& w, double t) {
  return t >= w.launch_time && !(t > w.launch_time + t_duration(w));
}

// launch_em_wall_2d / 3d, direction 1: rows y = 1..ny (and z = 1..nz, y fastest) of the
// column (the guard rows it also writes are replaced by the periodic refresh that
// follows).  second: the circular polarisation's second launch (phase0 + sign(pi/2,
// pol_type), pol + pi/2).
inline void column(const Wall& w0, int dims, int ny, int nz, const double* dx, float dtdx1,
                   int order, double t, bool second, float* b12, float* b3) {
  Wall w = w0;
  if (second) {
    w.phase0 = w.phase0 + (w.pol_type > 0 ? kPi2 : -kPi2);
    w.pol = w.pol + kPi2;
  }
  const double prop_sign = w.backward ? -1.0 : 1.0;
  double z = w.backward ? -w.xmax[0] : w.xmin[0];
  const double inj_time = t - w.launch_time;
  const double ldx = dx[0];
  const double rdtdx = (double)dtdx1;  // real(dt/dx, p_k_fld)
  const double ldxp = dx[1];
  double rmin = w.xmin[1];  // one node: nx_p_min = 1
  const double cos_pol = std::cos(w.pol) * prop_sign;
  const double sin_pol = -std::sin(w.pol) * prop_sign;
  const double amp = 2.0 * w.omega0 * w.a0;
  const double dr_2 = ldxp / 2.0;
  if (order == 2) {
    z = z + ldx * 0.5;
    rmin = rmin + dr_2;
  }
  const double lenv = amp * t_envelope(w, inj_time);
  if (dims == 2) {
    for (int ip = 1; ip <= ny; ip++) {
      const double r = rmin + (ip - 1) * ldxp;
      const double r_2 = r + dr_2;
      b12[ip - 1] = (float)(lenv * rdtdx * per_envelope_2d(w, z, r, inj_time) * sin_pol);
      b3[ip - 1] = (float)(lenv * rdtdx * per_envelope_2d(w, z, r_2, inj_time) * cos_pol);
    }
    return;
  }
  // 3D (:1015-1075): the lower corner rounded to p_k_fld as the reference does
  const double ldxp2 = dx[2], dr2_2 = ldxp2 / 2.0;
  double rmin1 = (double)(float)w.xmin[1], rmin2 = (double)(float)w.xmin[2];
  if (order == 2) {
    rmin1 = rmin1 + dr_2;
    rmin2 = rmin2 + dr2_2;
  }
  for (int ip2 = 1; ip2 <= nz; ip2++)
    for (int ip1 = 1; ip1 <= ny; ip1++) {
      const double r1 = rmin1 + (ip1 - 1) * ldxp;
      const double r2 = rmin2 + (ip2 - 1) * ldxp2;
      const double r1_2 = r1 + dr_2, r2_2 = r2 + dr2_2;
      const size_t o = (size_t)(ip2 - 1) * ny + (ip1 - 1);
      b12[o] = (float)(lenv * rdtdx * per_envelope_3d(w, z, r1, r2_2, inj_time) * sin_pol);
      b3[o] = (float)(lenv * rdtdx * per_envelope_3d(w, z, r1_2, r2, inj_time) * cos_pol);
    }
}

// read_input_zpulse_wall for direction 1 (:196-628): resolved parameters, or an error text
inline const char* resolve(const osr_zpulse_wall& p, int dims, const int* nx, const double* dx,
                           Wall& w) {
  w = Wall{};
  if (p.pol_type < -1 || p.pol_type > 1) return "pol_type must be -1, 0 or +1";
  if (p.propagation != OSR_PROP_FORWARD && p.propagation != OSR_PROP_BACKWARD)
    return "propagation must be forward or backward";
  w.a0 = p.a0;
  w.omega0 = p.omega0;
  w.phase0 = p.phase * kPi180;
  w.pol_type = p.pol_type;
  w.pol = p.pol * kPi180;
  w.backward = p.propagation == OSR_PROP_BACKWARD;
  for (int d = 0; d < dims; d++) {
    w.xmin[d] = p.xmin[d];
    w.xmax[d] = p.xmin[d] + nx[d] * dx[d];
  }
  w.tenv_type = p.tenv_type;
  switch (p.tenv_type) {
    case OSR_TENV_POLYNOMIAL:
    case OSR_TENV_SIN2:
      if (p.tenv_fwhm > 0.0) {
        w.rise = p.tenv_fwhm / 2.0;
        w.flat = 0.0;
        w.fall = p.tenv_fwhm / 2.0;
      } else {
        w.rise = p.tenv_rise;
        w.flat = p.tenv_flat;
        w.fall = p.tenv_fall;
      }
      break;
    case OSR_TENV_GAUSSIAN:
      w.duration = p.tenv_fwhm > 0.0 ? p.tenv_fwhm / std::sqrt(2.0 * std::log(2.0)) : p.tenv_duration;
      w.range = p.tenv_range;
      if (p.tenv_range == OSR_ZP_UNSET) return "a gaussian envelope needs tenv_range";
      break;
    default:
      return "tenv_type must be polynomial, sin2 or gaussian";
  }
  w.per_type = p.per_type;
  switch (p.per_type) {
    case OSR_PER_PLANE:
      break;
    case OSR_PER_HERMITE:
      for (int d = 0; d < dims - 1; d++) {
        if (p.per_tem_mode[d] < 0 || p.per_tem_mode[d] > 7) return "per_tem_mode must be 0..7";
        w.tem_mode[d] = p.per_tem_mode[d];
      }
      w.w0 = p.per_fwhm > 0.0 ? p.per_fwhm / std::sqrt(4.0 * std::log(2.0)) : p.per_w0;
      if (w.w0 <= 0.0) return "a gaussian beam needs per_w0 or per_fwhm > 0";
      if (p.per_focus == OSR_ZP_UNSET) return "a gaussian beam needs per_focus";
      w.focus = p.per_focus;
      break;
    default:
      return "per_type must be plane or hermite";
  }
  for (int d = 0; d < dims - 1; d++)  // :590-598
    w.center[d] = p.per_center[d] == OSR_ZP_UNSET ? 0.5 * (w.xmin[d + 1] + w.xmax[d + 1])
                                                   : p.per_center[d];
  w.launch_time = p.launch_time;
  if (w.backward) w.focus = -w.focus;
  return nullptr;
}

// ------------------------------------------------------------------------------------
// In-box pulse (zpulse, zpulse/os-zpulse-std.f03): launched once into the fields along x1,
// no boost, tilt or chirp, one node.  Box holds the parameters after read_input_zpulse
// (:157-805); the fields are built in the reference's guarded layout (f(3, 1-gcl:nx+gcu,
// ...), components interleaved) by launch_zpulse_2d / 3d (:1072-1387), in FP64 rounded
// to float, and corrected by div_corr_zpulse_2d / 3d (:1681-2205).
// ------------------------------------------------------------------------------------
struct Box {
  double a0, omega0, phase0, pol;
  int pol_type, backward, lon_type, per_type, tem_mode[2], div_corr;
  double lon_start, rise, flat, fall, lon_x0, duration, range;
  double center[2], w0, focus;
  double xmin[3];
};

// lon_envelope_zpulse :2480-2561 (the time is ignored without boost)
inline double box_lon(const Box& p, double z) {
  switch (p.lon_type) {
    case OSR_TENV_POLYNOMIAL:
      return p.backward ? fenv_poly(z - p.lon_start, p.rise, p.flat, p.fall)
                        : fenv_poly(p.lon_start - z, p.rise, p.flat, p.fall);
    case OSR_TENV_SIN2:
      return p.backward ? fenv_sin2(z - p.lon_start, p.rise, p.flat, p.fall)
                        : fenv_sin2(p.lon_start - z, p.rise, p.flat, p.fall);
    default: {
      if (std::fabs(z - p.lon_x0) > p.range / 2) return 0.0;
      const double a = (z - p.lon_x0) / p.duration;
      return std::exp(-2 * (a * a));
    }
  }
}

// lon_center_zpulse :3055-3097
inline double box_center(const Box& p) {
  if (p.lon_type == OSR_TENV_GAUSSIAN) return p.lon_x0;
  return p.backward ? p.lon_start + p.rise + p.flat / 2.0 : p.lon_start - p.rise - p.flat / 2.0;
}

// per_envelope_2d_zpulse / per_envelope_3d_zpulse at time 0 (plane, hermite_gaussian)
inline double box_per2(const Box& p, double phase0, double x1, double x2) {
  const double zc = box_center(p), k = p.omega0, ps = p.backward ? -1.0 : 1.0, t = 0.0;
  if (p.per_type == OSR_PER_PLANE) return std::cos(k * (x1 - zc - ps * t) + phase0);
  const double z0 = k * (p.w0 * p.w0) * 0.5;
  const double rho = x2 - p.center[0], rho2 = rho * rho, z = x1 - p.focus;
  double rWl2 = 1, curv = 0, gouy = 0;
  if (z != 0) {
    rWl2 = (z0 * z0) / ((z0 * z0) + (z * z));
    curv = 0.5 * rho2 * z / ((z * z) + (z0 * z0));
    gouy = (p.tem_mode[0] + 1) * std::atan2(z, z0);
  }
  const double rWu = std::sqrt(2 * rWl2) / p.w0;
  return std::sqrt(std::sqrt(rWl2)) * hermite(p.tem_mode[0], rho * rWu) *
         std::exp(-(rho2 * rWl2 / (p.w0 * p.w0))) *
         std::cos(k * (x1 - zc - ps * t) + k * curv - gouy + phase0);
}
inline double box_per3(const Box& p, double phase0, double x1, double x2, double x3) {
  const double zc = box_center(p), k = p.omega0, ps = p.backward ? -1.0 : 1.0, t = 0.0;
  const double ox = x2 - p.center[0], oy = x3 - p.center[1];
  if (p.per_type == OSR_PER_PLANE) return std::cos(k * (x1 - zc - ps * t) + phase0);
  const double rho2 = (ox * ox) + (oy * oy), z = x1 - p.focus, z0 = k * (p.w0 * p.w0) / 2;
  double rWl2 = 1, curv = 0, gouy = 0;
  if (z != 0) {
    rWl2 = (z0 * z0) / ((z0 * z0) + (z * z));
    curv = 0.5 * rho2 * z / ((z * z) + (z0 * z0));
    gouy = (p.tem_mode[0] + p.tem_mode[1] + 1) * std::atan2(z, z0);
  }
  const double rWu = std::sqrt(2 * rWl2) / p.w0;
  return std::sqrt(rWl2) * hermite(p.tem_mode[0], ox * rWu) * hermite(p.tem_mode[1], oy * rWu) *
         std::exp(-(rho2 * rWl2 / (p.w0 * p.w0))) *
         std::cos(k * (x1 - zc - ps * t) + k * curv - gouy + phase0);
}

// Guarded host grid of 3 interleaved components: nx interior cells, gcl / gcu guards.
struct GGrid {
  int dims, nx[3], gcl[3], gcu[3];
  long ld(int d) const { return d < dims ? (long)nx[d] + gcl[d] + gcu[d] : 1; }
  size_t cells() const { return (size_t)ld(0) * ld(1) * ld(2); }
  // component c of Fortran cell (i, j, k)
  size_t at(int c, int i, int j, int k) const {
    size_t o = (size_t)(i - 1 + gcl[0]) + (size_t)ld(0) * (size_t)(j - 1 + gcl[1]);
    if (dims == 3) o += (size_t)ld(0) * ld(1) * (size_t)(k - 1 + gcl[2]);
    return 3 * o + c;
  }
};

// One pulse (phase0, pol) into the zeroed arrays e, b (launch_zpulse_2d / 3d, then
// div_corr_zpulse_2d / 3d when p.div_corr).  order 2 shifts the origin by half a cell.
inline void box_build(const Box& p, const GGrid& g, int order, const double* dx, double phase0,
                      double pol, float* e, float* b) {
  const int D = g.dims;
  int lo[3] = {1, 1, 1}, hi[3] = {1, 1, 1};
  for (int d = 0; d < D; d++) {
    lo[d] = 1 - g.gcl[d];
    hi[d] = g.nx[d] + g.gcu[d];
  }
  const double cos_pol = std::cos(pol), sin_pol = std::sin(pol), amp = p.omega0 * p.a0;
  const double ps = p.backward ? -1.0 : 1.0;
  double zmin = p.xmin[0], rmin = p.xmin[1], r2min = D == 3 ? p.xmin[2] : 0.0;
  const double dz_2 = dx[0] / 2.0, dr_2 = dx[1] / 2.0, dr2_2 = D == 3 ? dx[2] / 2.0 : 0.0;
  if (order == 2) {
    zmin = zmin + dz_2;
    rmin = rmin + dr_2;
    r2min = r2min + dr2_2;
  }
  for (int i1 = lo[0]; i1 <= hi[0]; i1++) {
    const double z = zmin + (double)(i1 - 1) * dx[0], z_2 = z + dz_2;
    const double lenv = amp * box_lon(p, z), lenv_2 = amp * box_lon(p, z_2);
    if (D == 2) {  // :1127-1147, fields assigned
      for (int i2 = lo[1]; i2 <= hi[1]; i2++) {
        const double r = rmin + (double)(i2 - 1) * dx[1], r_2 = r + dr_2;
        b[g.at(0, i1, i2, 1)] = 0.0f;
        b[g.at(1, i1, i2, 1)] = (float)(-(lenv_2 * box_per2(p, phase0, z_2, r) * sin_pol * ps));
        b[g.at(2, i1, i2, 1)] = (float)(+(lenv_2 * box_per2(p, phase0, z_2, r_2) * cos_pol * ps));
        e[g.at(0, i1, i2, 1)] = 0.0f;
        e[g.at(1, i1, i2, 1)] = (float)(lenv * box_per2(p, phase0, z, r_2) * cos_pol);
        e[g.at(2, i1, i2, 1)] = (float)(lenv * box_per2(p, phase0, z, r) * sin_pol);
      }
      continue;
    }
    for (int i2 = lo[1]; i2 <= hi[1]; i2++) {  // :1288-1321, fields accumulated
      const double r1 = rmin + (double)(i2 - 1) * dx[1], r1_2 = r1 + dr_2;
      for (int i3 = lo[2]; i3 <= hi[2]; i3++) {
        const double r2 = r2min + (double)(i3 - 1) * dx[2], r2_2 = r2 + dr2_2;
        float& b2 = b[g.at(1, i1, i2, i3)];
        float& b3 = b[g.at(2, i1, i2, i3)];
        float& e2 = e[g.at(1, i1, i2, i3)];
        float& e3 = e[g.at(2, i1, i2, i3)];
        b2 = b2 - (float)(lenv_2 * box_per3(p, phase0, z_2, r1, r2_2) * sin_pol * ps);
        b3 = b3 + (float)(lenv_2 * box_per3(p, phase0, z_2, r1_2, r2) * cos_pol * ps);
        e2 = e2 + (float)(lenv * box_per3(p, phase0, z, r1_2, r2) * cos_pol);
        e3 = e3 + (float)(lenv * box_per3(p, phase0, z, r1, r2_2) * sin_pol);
      }
    }
  }
  if (!p.div_corr) return;
  // E1 and B1 integrated from the upper x edge, line by line over the interior (y, z);
  // one node: the other nodes' x range is empty, so e1 = b1 = 0 at x = nx + 1
  double x1min = p.xmin[0];
  if (order == 2) x1min = x1min + dx[0] * 0.5;
  const int n3 = D == 3 ? g.nx[2] : 1;
  for (int i3 = 1; i3 <= n3; i3++)
    for (int i2 = 1; i2 <= g.nx[1]; i2++) {
      double e1 = 0.0, b1 = 0.0;
      e[g.at(0, g.nx[0] + 1, i2, i3)] = (float)e1;
      b[g.at(0, g.nx[0] + 1, i2, i3)] = (float)b1;
      for (int i1 = g.nx[0]; i1 >= 1 - g.gcl[0]; i1--) {
        if (box_lon(p, x1min + i1 * dx[0]) > 0.0) {
          const double e2p = e[g.at(1, i1 + 1, i2, i3)], e2m = e[g.at(1, i1 + 1, i2 - 1, i3)];
          if (D == 2) {
            e1 = e1 + dx[0] * (e2p - e2m) / dx[1];
          } else {
            const double e3p = e[g.at(2, i1 + 1, i2, i3)], e3m = e[g.at(2, i1 + 1, i2, i3 - 1)];
            e1 = e1 + dx[0] * ((e2p - e2m) / dx[1] + (e3p - e3m) / dx[2]);
          }
          e[g.at(0, i1, i2, i3)] = (float)e1;
        } else {
          e1 = 0.0;
        }
        if (box_lon(p, x1min + (i1 - 0.5) * dx[0]) > 0.0) {
          const double b2p = b[g.at(1, i1, i2 + 1, i3)], b2m = b[g.at(1, i1, i2, i3)];
          if (D == 2) {
            b1 = b1 + dx[0] * (b2p - b2m) / dx[1];
          } else {
            const double b3p = b[g.at(2, i1, i2, i3 + 1)], b3m = b[g.at(2, i1, i2, i3)];
            b1 = b1 + dx[0] * ((b2p - b2m) / dx[1] + (b3p - b3m) / dx[2]);
          }
          b[g.at(0, i1, i2, i3)] = (float)b1;
        } else {
          b1 = 0.0;
        }
      }
    }
}

// read_input_zpulse for direction 1 (:246-805) and init_zpulse (:867-929): resolved
// parameters, or an error text
inline const char* resolve_box(const osr_zpulse& p, int dims, const int* nx, const double* dx,
                               Box& w) {
  w = Box{};
  if (p.pol_type < -1 || p.pol_type > 1) return "pol_type must be -1, 0 or +1";
  if (p.propagation != OSR_PROP_FORWARD && p.propagation != OSR_PROP_BACKWARD)
    return "propagation must be forward or backward";
  w.a0 = p.a0;
  w.omega0 = p.omega0;
  w.phase0 = p.phase * kPi180;
  w.pol_type = p.pol_type;
  w.pol = p.pol != 0.0 ? p.pol * kPi180 : 0.0;
  w.backward = p.propagation == OSR_PROP_BACKWARD;
  double xmax[3] = {0, 0, 0};
  for (int d = 0; d < dims; d++) {
    w.xmin[d] = p.xmin[d];
    xmax[d] = p.xmin[d] + nx[d] * dx[d];
  }
  w.lon_type = p.lon_type;
  switch (p.lon_type) {
    case OSR_TENV_POLYNOMIAL:
    case OSR_TENV_SIN2:
      if (p.lon_fwhm > 0.0) {
        w.rise = p.lon_fwhm / 2.0;
        w.flat = 0.0;
        w.fall = p.lon_fwhm / 2.0;
      } else {
        w.rise = p.lon_rise;
        w.flat = p.lon_flat;
        w.fall = p.lon_fall;
      }
      if (p.lon_start == OSR_ZP_UNSET) return "polynomial and sin2 envelopes need lon_start";
      w.lon_start = p.lon_start;
      break;
    case OSR_TENV_GAUSSIAN:
      w.duration = p.lon_fwhm > 0.0 ? p.lon_fwhm / std::sqrt(2.0 * std::log(2.0)) : p.lon_duration;
      if (p.lon_range == OSR_ZP_UNSET) return "a gaussian envelope needs lon_range";
      if (p.lon_x0 == OSR_ZP_UNSET) return "a gaussian envelope needs lon_x0";
      w.range = p.lon_range;
      w.lon_x0 = p.lon_x0;
      break;
    default:
      return "lon_type must be polynomial, sin2 or gaussian";
  }
  w.per_type = p.per_type;
  switch (p.per_type) {
    case OSR_PER_PLANE:
      break;
    case OSR_PER_HERMITE:
      for (int d = 0; d < dims - 1; d++) {
        if (p.per_tem_mode[d] < 0 || p.per_tem_mode[d] > 7) return "per_tem_mode must be 0..7";
        w.tem_mode[d] = p.per_tem_mode[d];
      }
      w.w0 = p.per_fwhm > 0.0 ? p.per_fwhm / std::sqrt(4.0 * std::log(2.0)) : p.per_w0;
      if (w.w0 <= 0.0) return "a gaussian beam needs per_w0 or per_fwhm > 0";
      if (p.per_focus == OSR_ZP_UNSET) return "a gaussian beam needs per_focus";
      w.focus = p.per_focus;
      break;
    default:
      return "per_type must be plane or hermite";
  }
  for (int d = 0; d < dims - 1; d++)  // :741-758
    w.center[d] = p.per_center[d] == OSR_ZP_UNSET ? 0.5 * (w.xmin[d + 1] + xmax[d + 1])
                                                   : p.per_center[d];
  w.div_corr = (w.per_type != OSR_PER_PLANE && !p.no_div_corr) ? 1 : 0;
  return nullptr;
}

}  // namespace osr_zp
