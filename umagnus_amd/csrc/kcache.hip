// Demag tensor cache: mag.DemagKernel / mag.NewellDemagKernel with a cache
// directory (mag/demagkernel.go:17-109, mag/newellDemagKernel.go:1006-1106).
//
// The tensor of a plan is looked up as one OVF file per component,
//   <dir>/uMagNUSkernel_<N>_<pbc>_<cellsize>_<accuracy>_<ij>.ovf      (brute force)
//   <dir>/uMagNUSkernel_<N>_<pbc>_<cellsize>_<zero_self_demag>_<ij>.ovf (Newell)
// with Go's fmt.Sprint spelling of the arrays ("[128 32 1]",
// "[3.90625e-09 3.90625e-09 1.5e-09]"), so the reference and this library
// share cache files.  ij runs over 00 01 02 11 12 22, without 02 and 12 when
// Nz == 1.  Files are OVF2 "binary 4" (float32) scalar fields of the padded
// size, titled N_ij, as SaveKernel writes them.
//
// Any failure to read (missing file, unknown header key, wrong size) falls
// back to computing the tensor on the GPU, which is then written to the cache
// (write failures are ignored), as the reference does.  With a cache
// directory the tensor used is always the float32 one the files hold (a fresh
// tensor is rounded before use), so a run does not depend on whether the
// cache was warm.  Without one (the default) the FP64 tensor is used.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>

#include "common.hpp"
#include "sim.hpp"

namespace um {

static std::mutex g_cache_mu;
static std::string g_cache_dir;

void set_kernel_cache(const char *dir) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  g_cache_dir = dir ? dir : "";
}
static std::string cache_dir() {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  return g_cache_dir;
}

// Go's %v of a float64: the shortest decimal that reads back to v (strconv
// 'g' with precision -1), in exponent form when the decimal exponent is below
// -4 or at least 6, with at least two exponent digits.
std::string go_float(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "+Inf" : "-Inf";
  if (v == 0) return std::signbit(v) ? "-0" : "0";
  char buf[40];
  int prec = 0;
  for (prec = 0; prec < 17; prec++) {   // fewest significant digits that round-trip
    std::snprintf(buf, sizeof buf, "%.*e", prec, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  // buf = [-]d[.ddd]e(+|-)xx
  std::string s(buf);
  const size_t e = s.find('e');
  const int exp = std::atoi(s.c_str() + e + 1);
  std::string mant = s.substr(0, e);
  const bool neg = mant[0] == '-';
  if (neg) mant.erase(0, 1);
  std::string digits;
  for (char ch : mant)
    if (ch != '.') digits += ch;
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string out = neg ? "-" : "";
  if (exp < -4 || exp >= 6) {
    out += digits[0];
    if (digits.size() > 1) out += "." + digits.substr(1);
    char eb[8];
    std::snprintf(eb, sizeof eb, "e%c%02d", exp < 0 ? '-' : '+', exp < 0 ? -exp : exp);
    out += eb;
  } else if (exp >= 0) {
    std::string ip = digits.substr(0, std::min<size_t>(digits.size(), exp + 1));
    while ((int)ip.size() < exp + 1) ip += '0';
    out += ip;
    if ((int)digits.size() > exp + 1) out += "." + digits.substr(exp + 1);
  } else {
    out += "0." + std::string(-exp - 1, '0') + digits;
  }
  return out;
}

// fmt.Sprint of [3]int / [3]float64
static std::string go_ints(const int v[3]) {
  return "[" + std::to_string(v[0]) + " " + std::to_string(v[1]) + " " + std::to_string(v[2]) + "]";
}
static std::string go_floats(const double v[3]) {
  return "[" + go_float(v[0]) + " " + go_float(v[1]) + " " + go_float(v[2]) + "]";
}

// basename of mag/demagkernel.go:40 (accuracy) and newellDemagKernel.go:1023
// (zero_self_demag; the asymptotic radius is not part of the key)
std::string kernel_cache_basename(const std::string &dir, const int N[3], const int pbc[3], const double cs[3],
                                  int newell, double accuracy, int zero_self_demag) {
  const std::string tail = newell ? std::to_string(zero_self_demag) : go_float(accuracy);
  return dir + "/" + "uMagNUSkernel_" + go_ints(N) + "_" + go_ints(pbc) + "_" + go_floats(cs) + "_" + tail + "_";
}

// ---- OVF (oommf/oommf.go:17-214, ovf1.go, ovf2.go) ----------------------------
static std::string lower(std::string s) {
  for (auto &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}
static std::string trim(const std::string &s, const char *set) {
  const size_t a = s.find_first_not_of(set);
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(set);
  return s.substr(a, b - a + 1);
}
// parseHeaderLine (oommf.go:172-181)
static void header_line(const std::string &line, std::string &key, std::string &value) {
  const size_t c = line.find(':');
  key = trim(line.substr(0, c), "# ");
  value = c == std::string::npos ? "" : trim(line.substr(c + 1), "# ");
}
static bool read_line(std::istream &in, std::string &line) {
  if (!std::getline(in, line)) return false;
  if (!line.empty() && line.back() == '\r') line.pop_back();
  return true;
}
static long to_long(const std::string &s) {
  char *end = nullptr;
  const long v = std::strtol(s.c_str(), &end, 10);
  UM_REQUIRE(end && end != s.c_str() && trim(end, " \t").empty(), UM_EINVAL, "OVF: not an integer: " + s);
  return v;
}

// ReadFile (oommf.go:17-58): a scalar or vector field, values in FP64 as stored.
void ovf_read(const std::string &fname, int size[3], int *ncomp, std::vector<double> &vals) {
  std::ifstream in(fname, std::ios::binary);
  UM_REQUIRE(in.good(), UM_EINVAL, "OVF: cannot open " + fname);
  std::string line;
  UM_REQUIRE(read_line(in, line), UM_EINVAL, "OVF: empty file");
  int ovf = 0;
  *ncomp = 0;
  const std::string tag = lower(line);
  if (tag == "# oommf ovf 2.0") ovf = 2;
  else if (tag == "# oommf: rectangular mesh v1.0") { ovf = 1; *ncomp = 3; }
  else throw Error(UM_EINVAL, "OVF: unknown header: " + line);
  size[0] = size[1] = size[2] = 0;
  std::string key, value;
  for (;;) {
    UM_REQUIRE(read_line(in, line), UM_EINVAL, "OVF: missing Begin: Data");
    std::string h = lower(trim(line, "# "));
    h.erase(std::remove(h.begin(), h.end(), ' '), h.end());
    if (h.rfind("begin:data", 0) == 0) break;   // isHeaderEnd
    header_line(line, key, value);
    const std::string k = lower(key);
    static const char *ignored[] = {"oommf", "segment count", "begin", "meshtype", "xbase", "ybase", "zbase",
                                    "xmin", "ymin", "zmin", "xmax", "ymax", "zmax", "valuerangeminmag",
                                    "valuerangemaxmag", "end", "", "valuelabels", "title", "valueunits",
                                    "xstepsize", "ystepsize", "zstepsize", "valuemultiplier", "valueunit",
                                    "meshunit", "desc"};
    if (k == "valuedim") *ncomp = (int)to_long(value);
    else if (k == "xnodes") size[0] = (int)to_long(value);
    else if (k == "ynodes") size[1] = (int)to_long(value);
    else if (k == "znodes") size[2] = (int)to_long(value);
    else {
      bool known = false;
      for (const char *g : ignored) known = known || k == g;
      UM_REQUIRE(known, UM_EINVAL, "OVF: Unknown key: " + key);
    }
  }
  header_line(line, key, value);
  std::istringstream ws(trim(value, " \t"));
  std::string w0, w1, w2;
  ws >> w0 >> w1 >> w2;
  UM_REQUIRE(lower(key) == "begin" && lower(w0) == "data", UM_EINVAL, "OVF: Expected: Begin: Data");
  const std::string fmt = w2.empty() ? "text" : "binary " + w2;
  UM_REQUIRE(*ncomp > 0 && size[0] > 0 && size[1] > 0 && size[2] > 0, UM_EINVAL, "OVF: bad size");
  const size_t count = (size_t)*ncomp * size[0] * size[1] * size[2];
  std::vector<double> flat(count);
  if (fmt == "text") {
    for (size_t i = 0; i < count;) {
      UM_REQUIRE(read_line(in, line), UM_EINVAL, "OVF: unexpected end of text data");
      if (!line.empty() && line[0] == '#') continue;
      std::istringstream ls(line);
      double v;
      while (i < count && (ls >> v)) flat[i++] = v;
    }
  } else if (fmt == "binary 4" || fmt == "binary 8") {
    const int wb = fmt == "binary 4" ? 4 : 8;
    auto rd = [&](double *out) {
      unsigned char b[8];
      in.read((char *)b, wb);
      UM_REQUIRE(in.gcount() == wb, UM_EINVAL, "OVF: unexpected end of binary data");
      if (ovf == 1) std::reverse(b, b + wb);   // OVF1 is big endian, OVF2 little endian (as the host)
      if (wb == 4) { float f; std::memcpy(&f, b, 4); *out = f; }
      else std::memcpy(out, b, 8);
    };
    double ctrl;
    rd(&ctrl);
    UM_REQUIRE(ctrl == (wb == 4 ? 1234567.0 : 123456789012345.0), UM_EINVAL, "OVF: invalid control number");
    for (size_t i = 0; i < count; i++) rd(&flat[i]);
  } else {
    throw Error(UM_EINVAL, "OVF: unknown format: " + fmt);
  }
  // file order: x fastest, components innermost -> SoA
  const size_t nc = (size_t)size[0] * size[1] * size[2];
  vals.resize(count);
  for (size_t i = 0; i < nc; i++)
    for (int c = 0; c < *ncomp; c++) vals[c * nc + i] = flat[i * *ncomp + c];
}

// SaveKernel (mag/demagkernel.go:102-109) = WriteOVF2(..., "binary 4") of a
// scalar field titled `name` (oommf/ovf2.go:13-80, oommf.go:203-208).
void ovf_write_scalar(const std::string &fname, const int size[3], const double cs[3], const std::string &name,
                      const double *K) {
  std::ostringstream h;
  auto hdr = [&](const std::string &key, const std::string &value) { h << "# " << key << ": " << value << "\n"; };
  h << "# OOMMF OVF 2.0\n";
  hdr("Segment count", "1");
  hdr("Begin", "Segment");
  hdr("Begin", "Header");
  hdr("Title", name);
  hdr("meshtype", "rectangular");
  hdr("meshunit", "m");
  hdr("xmin", "0");
  hdr("ymin", "0");
  hdr("zmin", "0");
  hdr("xmax", go_float(cs[0] * size[0]));
  hdr("ymax", go_float(cs[1] * size[1]));
  hdr("zmax", go_float(cs[2] * size[2]));
  hdr("valuedim", "1");
  hdr("valuelabels", name);
  hdr("valueunits", "1");
  hdr("Desc", "Total simulation time:  0  s");   // Fprintln puts blanks between operands
  hdr("xbase", go_float(cs[0] / 2));
  hdr("ybase", go_float(cs[1] / 2));
  hdr("zbase", go_float(cs[2] / 2));
  hdr("xnodes", std::to_string(size[0]));
  hdr("ynodes", std::to_string(size[1]));
  hdr("znodes", std::to_string(size[2]));
  hdr("xstepsize", go_float(cs[0]));
  hdr("ystepsize", go_float(cs[1]));
  hdr("zstepsize", go_float(cs[2]));
  hdr("End", "Header");
  hdr("Begin", "Data Binary 4");
  const size_t n = (size_t)size[0] * size[1] * size[2];
  std::vector<float> buf(n + 1);
  buf[0] = 1234567.0f;
  for (size_t i = 0; i < n; i++) buf[i + 1] = (float)K[i];
  std::ofstream out(fname, std::ios::binary | std::ios::trunc);
  UM_REQUIRE(out.good(), UM_EINVAL, "OVF: cannot create " + fname);
  const std::string hs = h.str();
  out.write(hs.data(), (std::streamsize)hs.size());
  out.write((const char *)buf.data(), (std::streamsize)(sizeof(float) * buf.size()));
  out << "# End: Data Binary 4\n# End: Segment\n";   // directly after the data, as hdr() writes it
  UM_REQUIRE(out.good(), UM_EINVAL, "OVF: write failed: " + fname);
}

// component ij of the reference -> index in xx yy zz yz xz xy
static const int kIJ[6][3] = {{0, 0, 0}, {0, 1, 5}, {0, 2, 4}, {1, 1, 1}, {1, 2, 3}, {2, 2, 2}};

static __global__ void k_round_f32(double *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = (double)(float)p[i];
}

// The tensor of a plan into Kdev[6] (FP64, padded size P): from the cache when
// a directory is set and every component reads back, else computed on the GPU.
void demag_tensor(const int N[3], const int pbc[3], const double cs[3], int newell, double accuracy,
                  int asymptotic_radius, int zero_self_demag, double *const Kdev[6], int P[3], hipStream_t s) {
  // sanityCheck (mag/demagkernel.go:385-396): cell aspect ratios within 100
  const double a3 = cs[0] / cs[1], a2 = cs[1] / cs[2], a1 = cs[2] / cs[0];
  const double amax = std::max(a1, std::max(a2, a3)), amin = std::min(a1, std::min(a2, a3));
  UM_REQUIRE(!(amax > 100.0 || amin < 1.0 / 100.0), UM_EINVAL, "Unrealistic cell aspect ratio");
  const std::string dir = cache_dir();
  pad_size(N, pbc, P);
  const size_t np = (size_t)P[0] * P[1] * P[2];
  const bool is2d = N[2] == 1;
  if (!dir.empty()) {
    const std::string base = kernel_cache_basename(dir, N, pbc, cs, newell, accuracy, zero_self_demag);
    try {
      std::vector<double> host[6];
      for (auto &ij : kIJ) {
        if (is2d && ij[1] == 2 && ij[0] != 2) continue;   // xz, yz not stored in 2D
        int sz[3], nc;
        ovf_read(base + std::to_string(ij[0]) + std::to_string(ij[1]) + ".ovf", sz, &nc, host[ij[2]]);
        UM_REQUIRE(nc == 1 && sz[0] == P[0] && sz[1] == P[1] && sz[2] == P[2], UM_EINVAL, "cached kernel size");
      }
      for (int c = 0; c < 6; c++) {
        if (host[c].empty()) fill_value<double>(Kdev[c], 0.0, np, s);
        else UM_HIP(hipMemcpyAsync(Kdev[c], host[c].data(), sizeof(double) * np, hipMemcpyHostToDevice, s));
      }
      UM_HIP(hipStreamSynchronize(s));
      return;
    } catch (const Error &) {
      // not cached (or unreadable): compute and store below
    }
  }
  int P2[3];
  if (newell) demag_kernel_newell_device(N, pbc, cs, asymptotic_radius, zero_self_demag, Kdev, P2, s);
  else demag_kernel_bf_device(N, pbc, cs, accuracy, Kdev, P2, s);
  if (dir.empty()) return;
  for (int c = 0; c < 6; c++) {
    UM_LAUNCH(k_round_f32, dim3(blocks_for((long)np, 256, 4096)), dim3(256), 0, s, Kdev[c], np);
    UM_HIP(hipGetLastError());
  }
  const std::string base = kernel_cache_basename(dir, N, pbc, cs, newell, accuracy, zero_self_demag);
  std::vector<double> h(np);
  for (auto &ij : kIJ) {
    if (is2d && ij[1] == 2 && ij[0] != 2) continue;
    UM_HIP(hipMemcpyAsync(h.data(), Kdev[ij[2]], sizeof(double) * np, hipMemcpyDeviceToHost, s));
    UM_HIP(hipStreamSynchronize(s));
    try {
      ovf_write_scalar(base + std::to_string(ij[0]) + std::to_string(ij[1]) + ".ovf", P, cs,
                       "N_" + std::to_string(ij[0]) + std::to_string(ij[1]), h.data());
    } catch (const Error &) {
      break;   // "Failed to cache kernel": the computed tensor is used anyway
    }
  }
}

}  // namespace um
